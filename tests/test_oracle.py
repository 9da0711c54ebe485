"""Pins of the CPU oracle (no GPU): the reference's own known-answer tests,
hand-derived vectors and the mathematics each oracle function must obey.

Every check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign, a transposed index) fails it; none retypes the oracle's
formula.  Sources of the known answers:

* unpack: the VDIF 4-level table and bit order (format_vdif.c:66-96, byte b
  sample i = bits 2i+1:2i) decoded through an independently built byte table;
* FFT: brute-force FP64 DFT, numpy.fft, a single-tone spectrum, Parseval,
  linearity;
* MAC: minitest-cell.cpp:131-161 (acc += a*b with a=(0.5i,-0.5i),
  b=(0.15i^2,-0.25), acc0=(1,1.5)) and the conjugate closed form;
* rotation: rotate(tau) then rotate(-tau) is the identity; a delayed copy is
  re-aligned (cross phase flat across channels);
* X-engine: numpy products, Hermitian symmetry, real non-negative autos and
  the 2-bit Van Vleck ratio;
* phase-cal: PCal.cpp:879-951,969-980 (tones at -90 + 5 t degrees within 0.1
  degree, with the adjustSampleOffset skip path) for all three extractors;
* spectrometer: swspectrometer.m:7-48 (carrier 1/37 of the sample rate peaks
  at bin round(N/37)), mean-power normalisation and the window closed forms.
"""
import math

import numpy as np
import pytest

import oracle as O

H = 3.3359


def nrel(a, b):
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---------------------------------------------------------------- a1 unpack
def test_unpack_hand_vectors():
    # 0xE4 = 11 10 01 00: samples (bits 1:0, 3:2, 5:4, 7:6) are codes 0,1,2,3
    got = O.unpack_1ch(np.array([0xE4E4E4E4], np.uint32), 16)
    want = np.tile(np.array([-H, -1, 1, H], np.float32), 4)
    assert np.array_equal(got, want)
    got = O.unpack_1ch(np.array([0x1B1B1B1B, 0x00000000, 0xFFFFFFFF], np.uint32), 48)
    want = np.concatenate([np.tile(np.array([H, 1, -1, -H], np.float32), 4),
                           np.full(16, -H, np.float32), np.full(16, H, np.float32)])
    assert np.array_equal(got, want)
    # the constants themselves are the float32 of the table values
    assert np.array_equal(O.LUT, np.array([-3.3359, -1.0, 1.0, 3.3359], np.float32))


def _byte_table():
    """lut2bit[b][i] = level[(b >> 2i) & 3] built byte-wise as mark5access
    initialises it (format_vdif.c:83-89)."""
    level = np.array([-H, -1.0, 1.0, H], np.float32)
    return np.array([[level[(b >> (2 * i)) & 3] for i in range(4)] for b in range(256)], np.float32)


def test_unpack_matches_byte_table():
    rng = np.random.default_rng(5)
    words = rng.integers(0, 2**32, 257, dtype=np.uint64).astype(np.uint32)
    bytes_ = words.view(np.uint8)          # little-endian byte order of the words
    want = _byte_table()[bytes_].ravel()
    got = O.unpack_1ch(words, words.size * 16)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("nchan", [2, 4, 8, 16])
def test_unpack_multichannel_layout(nchan):
    # channel c of time sample t sits at field t*nchan + c of the word stream
    spw = 16 // nchan
    ntime = spw * 37
    codes = np.zeros(ntime * nchan, np.uint8)
    rng = np.random.default_rng(nchan)
    codes[:] = rng.integers(0, 4, codes.size)
    words = O.pack2b(codes)
    got = O.unpack_nch(words, ntime, nchan)
    want = O.LUT[codes.reshape(ntime, nchan).T]
    assert np.array_equal(got, want)
    # the 16-channel mark5b-style word: channel c carries code c % 4 everywhere
    if nchan == 16:
        w = sum((c % 4) << (2 * c) for c in range(16))
        got = O.unpack_nch(np.array([w] * 3, np.uint32), 3, 16)
        assert np.array_equal(got[:, 0], O.LUT[np.arange(16) % 4])


def test_pack_roundtrip_and_quantizer():
    codes = np.arange(64, dtype=np.uint8) % 4
    assert np.array_equal(O.unpack_1ch(O.pack2b(codes), 64), O.LUT[codes])
    q = O.load().fxo_quantize2
    assert [q(x, 0.9816) for x in (-2.0, -0.5, 0.0, 0.5, 0.9816, 3.0)] == [0, 1, 2, 2, 3, 3]


def test_generator_statistics():
    # 2-bit codes at thresholds +-0.9816 sigma: P = {0.163, 0.337, 0.337, 0.163}
    w = O.gen_station(7, 0, 0.0, 0, 0.0, 0.0, 0.9816, 1 << 18)
    x = O.unpack_1ch(w, 1 << 18)
    p = np.array([np.mean(x == v) for v in O.LUT])
    assert np.allclose(p, [0.1633, 0.3367, 0.3367, 0.1633], atol=0.005)


# ---------------------------------------------------------------- a2/a3 FFT
@pytest.mark.parametrize("n", [4, 8, 16, 64, 256, 1024])
def test_twiddles_closed_form(n):
    W = O.twiddles(n)
    i = np.arange(n // 4)
    want = np.cos(2 * np.pi * i / n) - 1j * np.sin(2 * np.pi * i / n)
    assert np.max(np.abs(W - want)) < 2e-7


@pytest.mark.parametrize("n", [2, 4, 8, 32, 128, 512, 2048])
def test_fft_vs_bruteforce_dft(n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    ref = O.dft_f64(x)
    # the brute-force DFT itself against its definition at a few bins
    k = np.arange(n)
    for kk in (0, 1, n // 2, n - 1):
        assert abs(ref[kk] - np.sum(x.astype(np.complex128) * np.exp(-2j * np.pi * kk * k / n))) < 1e-9 * n
    assert nrel(O.fft_f64(x), ref) < 1e-13
    assert nrel(O.fft_r2(x), ref) < 2e-6 * math.log2(max(n, 2))
    assert nrel(np.fft.fft(x.astype(np.complex128)), ref) < 1e-13


@pytest.mark.parametrize("n", [1024, 8192, 32768])
def test_rfft_vs_numpy(n):
    rng = np.random.default_rng(n + 1)
    x = O.LUT[rng.integers(0, 4, n)]
    got = O.rfft_r2(x)
    ref = np.fft.rfft(x.astype(np.float64))[: n // 2]
    assert nrel(got, ref) < 1e-6


def test_fft_single_tone_and_parseval_linearity():
    n = 1024
    t = np.arange(n)
    x = np.cos(2 * np.pi * 37 * t / n).astype(np.float32)
    X = O.rfft_r2(x)
    assert np.argmax(np.abs(X)) == 37
    assert abs(abs(X[37]) - n / 2) < 1e-3
    mask = np.ones(n // 2, bool)
    mask[37] = False
    assert np.max(np.abs(X[mask])) < 1e-2
    rng = np.random.default_rng(3)
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    b = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    A, B = O.fft_r2(a), O.fft_r2(b)
    assert nrel(O.fft_r2((2 * a + 3j * b).astype(np.complex64)), 2 * A + 3j * B) < 1e-6
    e_t = np.sum(np.abs(a.astype(np.complex128)) ** 2)
    e_f = np.sum(np.abs(A.astype(np.complex128)) ** 2) / n
    assert abs(e_t - e_f) / e_t < 1e-6


# ---------------------------------------------------------------- MAC KATs
def test_mac_known_answer():
    n = 64
    i = np.arange(n, dtype=np.float64)
    a = (0.5 * i - 0.5j * i).astype(np.complex64)
    b = (0.15 * i * i - 0.25j).astype(np.complex64)
    acc0 = np.full(n, 1.0 + 1.5j, np.complex64)
    got = O.addproduct(a, b, acc0, conj=False)
    want = (1 + 0.075 * i**3 - 0.125 * i) + 1j * (1.5 - 0.075 * i**3 - 0.125 * i)
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1)) < 1e-6
    # conjugate variant: acc += a * conj(b)
    gotc = O.addproduct(a, b, acc0, conj=True)
    wantc = (1 + 0.075 * i**3 + 0.125 * i) + 1j * (1.5 - 0.075 * i**3 + 0.125 * i)
    assert np.max(np.abs(gotc - wantc) / np.maximum(np.abs(wantc), 1)) < 1e-6
    # the GENERIC header's own four-term example, (1+2i)(3+4i)... as probed
    assert np.allclose(O.mul(np.array([1 + 2j], np.complex64), np.array([3 + 4j], np.complex64)), [-5 + 10j])


# ---------------------------------------------------------------- a4 rotation
def test_phase_reduction_and_inverse():
    m = (12.375, 1e-3, 0.3, 0.01)
    for k in (0, 1, 511, 4095):
        ph = O.phase(m, k, 8192, 17)
        assert -math.pi <= ph <= math.pi
        tau = 12.375 + 1e-3 * 17
        cyc = k * tau / 8192 + (0.3 + 0.01 * 17) / (2 * math.pi)
        assert abs(math.remainder(ph / (2 * math.pi) - cyc, 1.0)) < 1e-12
    rng = np.random.default_rng(4)
    X = (rng.standard_normal(4096) + 1j * rng.standard_normal(4096)).astype(np.complex64)
    Y = O.rotate(O.rotate(X, 8192, m, 17), 8192, tuple(-v for v in m), 17)
    assert nrel(Y, X) < 1e-6


def test_delay_compensation_flattens_cross_phase():
    n, ntime = 1024, 1024 * 256
    d = 5
    s0 = O.gen_station(21, 0, 0.6, 0, 0.0, 0.0, 0.9816, ntime)
    s1 = O.gen_station(21, 1, 0.6, d, 0.0, 0.0, 0.9816, ntime)
    k = np.arange(1, n // 2)
    raw = O.corr_f32([s0, s1], n, ntime // n)[1][1:]
    slope = np.polyfit(k, np.unwrap(np.angle(raw)), 1)[0]
    # station 1 lags by d samples: cross phase slope +2 pi d / n per channel
    assert abs(slope - 2 * np.pi * d / n) < 0.05 * 2 * np.pi * d / n
    models = [(0.0, 0.0, 0.0, 0.0), (float(d), 0.0, 0.0, 0.0)]
    fixed = O.corr_f32([s0, s1], n, ntime // n, models)[1][1:]
    assert abs(np.polyfit(k, np.unwrap(np.angle(fixed)), 1)[0]) < 2e-4
    assert abs(np.angle(np.sum(fixed))) < 0.05


# ---------------------------------------------------------------- a6 X-engine
def test_xengine_products_and_symmetry():
    rng = np.random.default_rng(9)
    Cn, T, I = 8, 32, 5
    X = (rng.standard_normal((Cn, T, I)) + 1j * rng.standard_normal((Cn, T, I))).astype(np.complex64)
    vis = O.xengine_f64(X)
    p = 0
    for i in range(I):
        for j in range(i, I):
            assert O.prod_index(i, j, I) == p
            want = np.sum(X[:, :, i].astype(np.complex128) * np.conj(X[:, :, j]), axis=1)
            assert nrel(vis[p], want) < 1e-12
            if i == j:
                assert np.all(np.abs(vis[p].imag) == 0) and np.all(vis[p].real >= 0)
            p += 1
    assert nrel(O.xengine_f32(X), vis) < 1e-6
    # swapping the two inputs conjugates the product
    Xs = X[:, :, ::-1].copy()
    v2 = O.xengine_f64(Xs)
    assert nrel(v2[O.prod_index(0, I - 1, I)], np.conj(vis[O.prod_index(0, I - 1, I)])) < 1e-12


def test_van_vleck_two_bit():
    # 2-bit quantized correlation is ~0.88 rho at the optimal thresholds;
    # two stations sharing a common part of weight a correlate at rho = a^2
    n, nt = 256, 8192
    rho = 0.1
    s = [O.gen_station(33, k, math.sqrt(rho), 0, 0.0, 0.0, 0.9816, n * nt) for k in range(2)]
    v = O.corr_f64(s, n, nt, nthreads=4)
    r = np.sum(v[1].real) / math.sqrt(np.sum(v[0].real) * np.sum(v[2].real))
    assert abs(r / rho - 0.88) < 0.03


def test_corr_f32_matches_f64():
    n, nt = 512, 64
    s = [O.gen_station(2, k, 0.05, k, 0.0, 0.0, 0.9816, n * nt) for k in range(3)]
    m = [(0.5 * k, 0.0, 0.1 * k, 0.0) for k in range(3)]
    assert nrel(O.corr_f32(s, n, nt, m, nthreads=2), O.corr_f64(s, n, nt, m)) < 1e-5


# ---------------------------------------------------------------- a5 phase-cal
def _pcal_kat(bw, spacing, offset, sampleoffset, nsamp, skip=True):
    cfg = O.pcal_config(bw, spacing, offset)
    nt = cfg.ntones
    n = np.arange(nsamp, dtype=np.float64)
    data = np.zeros(nsamp)
    for t in range(nt):
        data += np.sin(np.pi * (n + sampleoffset) * (offset + t * spacing) / bw + t * np.pi * 5 / 180)
    data = data.astype(np.float32)
    acc = O.pcal_extract(cfg, data, sampleoffset)
    if skip:
        # PCal.cpp:915-920: skip 3 samples, continue at the matching offset
        acc = O.pcal_extract(cfg, data[3:], sampleoffset + nsamp + 3, acc)
    tones = O.pcal_final(cfg, acc)
    phi = np.degrees(np.angle(tones))
    want = -90.0 + 5.0 * np.arange(nt)
    return cfg, tones, np.abs(((phi - want) + 180) % 360 - 180)


@pytest.mark.parametrize("bw,spacing,offset,kind", [
    (16e6, 1e6, 510e3, O.PCAL_IMPLICIT),      # PCal.cpp:868 example invocation
    (8e6, 1e6, 0, O.PCAL_TRIVIAL),            # PhaseCal/test.cpp:34
    (8e6, 1e6, 10e3, O.PCAL_IMPLICIT),        # PhaseCal/test.cpp:36 (1600-sample period)
    (8e6, 1e6, 400e3, O.PCAL_SHIFTING),       # period 40 not a multiple of 16 bins
])
def test_pcal_reference_kat(bw, spacing, offset, kind):
    cfg, tones, err = _pcal_kat(bw, spacing, offset, 0, 32000)
    assert cfg.kind == kind
    if offset == 0:
        # tone 0 is DC, where the KAT's sin(0) carries no signal (only the
        # leakage of the skipped partial periods)
        assert abs(tones[0]) < 1e-3 * np.max(np.abs(tones))
        err = err[1:]
    assert np.all(err < 0.1), err


def test_pcal_config_factory():
    c = O.pcal_config(16e6, 1e6, 510e3)
    assert (c.kind, c.nbins, c.ntones) == (O.PCAL_IMPLICIT, 3200, 16)
    c = O.pcal_config(8e6, 1e6, 10e3)
    assert (c.kind, c.nbins) == (O.PCAL_IMPLICIT, 1600)
    c = O.pcal_config(8e6, 1e6, 0)
    assert (c.kind, c.nbins, c.ntones) == (O.PCAL_TRIVIAL, 16, 8)
    c = O.pcal_config(8e6, 1e6, 400e3)
    assert (c.kind, c.nbins, c.rotatorlen) == (O.PCAL_SHIFTING, 16, 40)


def test_pcal_fast_equals_direct_method():
    # extractAndIntegrate vs the direct (quasi-reference) method of
    # PCal.cpp:139-175 on noise plus tones
    rng = np.random.default_rng(8)
    for bw, sp, off in ((8e6, 1e6, 10e3), (8e6, 1e6, 400e3), (8e6, 1e6, 0)):
        cfg = O.pcal_config(bw, sp, off)
        x = O.LUT[rng.integers(0, 4, 16000)]
        fast = O.pcal_final(cfg, O.pcal_extract(cfg, x, 7))
        direct = O.pcal_reference(cfg, x, 7)
        # each method scales by its own DFT length: nbins vs 2*ntones
        assert nrel(fast * (cfg.nbins / (2 * cfg.ntones)), direct) < 1e-9


def test_taskcore_pcal_fold_tones():
    # the IPP core's folded pulse: its tonebins-point DFT gives the tone
    # phasors (tone t at bin t for a zero-offset-aligned rotator)
    fs, spacing, offset = 16e6, 1e6, 500e3
    T = int(fs / spacing)
    R = int(fs / math.gcd(int(fs), int(offset)))
    assert (T, R) == (16, 32)
    npts = 1024
    n = np.arange(npts)
    x = np.zeros(npts)
    for t in range(T // 2):
        x += np.sin(2 * np.pi * n * (offset + t * spacing) / fs + t * np.pi * 5 / 180)
    acc = O.tcpcal_fold(x.astype(np.float32), T, R, -2 * np.pi * offset / fs)
    spec = np.fft.fft(acc) / (npts)
    phi = np.degrees(np.angle(spec[: T // 2]))
    want = -90.0 + 5.0 * np.arange(T // 2)
    assert np.all(np.abs(((phi - want) + 180) % 360 - 180) < 0.1)


# ---------------------------------------------------------------- a8 spectrometer
def test_window_closed_forms():
    n = 256
    i = np.arange(n)
    assert np.max(np.abs(O.window(O.WIN_HANN, n) - np.hanning(n))) < 1e-7
    assert np.max(np.abs(O.window(O.WIN_HAMMING, n) - np.hamming(n))) < 1e-7
    assert np.max(np.abs(O.window_blackman(-0.16, n) - np.blackman(n))) < 1e-7
    assert np.max(np.abs(O.window(O.WIN_NONE, n) - 1)) == 0
    # SIN2 is symmetric, zero at both ends, one in the middle of an odd grid
    w = O.window(O.WIN_SIN2, 257)
    assert w[0] < 1e-12 and w[-1] < 1e-12 and abs(w[128] - 1) < 1e-7
    # the cosine window peaks between the two middle samples
    wc = O.window(O.WIN_COSINE, n)
    assert np.allclose(wc, wc[::-1], atol=1e-7) and abs(wc[0] - math.cos(math.pi * (0.5 - n / 2) / n)) < 1e-7
    assert np.allclose(O.window(O.WIN_COSINE2, n), wc.astype(np.float64) ** 2, atol=1e-6)


def test_spectrometer_carrier_kat():
    # swspectrometer.m: carrier 1/37 in noise, 1024-point cos-windowed FFTs
    # with 50 % overlap; the power peak sits at round(1024/37) = 28
    n = 1024
    ntime = 1 << 18
    w = O.gen_multichannel(37, 1, [0.1], [1 / 37], 0.9816, ntime)
    out = O.spectrometer_ex_f32(w, ntime, 1, n, 2, 511, O.WIN_COSINE, True, True, nthreads=4)
    assert out.shape == (1, 1, n // 2 + 1)
    assert int(np.argmax(out[0, 0, 1:])) + 1 == 28


def test_spectrometer_mean_power_and_nyquist():
    n, K = 256, 8
    ntime = n * K * 2
    rng = np.random.default_rng(12)
    codes = rng.integers(0, 4, ntime).astype(np.uint8)
    w = O.pack2b(codes)
    out = O.spectrometer_ex_f32(w, ntime, 1, n, 1, K, O.WIN_NONE, False, True)
    x = O.LUT[codes].astype(np.float64).reshape(2, K, n)
    P = np.mean(np.abs(np.fft.rfft(x, axis=2)) ** 2, axis=1)
    assert nrel(out[:, 0, :], P) < 1e-6
    base = O.spectrometer_f32(w, ntime, 1, n, K, O.WIN_NONE)
    assert np.array_equal(base[:, 0, :], out[:, 0, : n // 2])


def test_spectrometer_overlap_conventions():
    n, K, ov = 256, 5, 2
    ntime = 4096
    rng = np.random.default_rng(13)
    codes = rng.integers(0, 4, ntime).astype(np.uint8)
    w = O.pack2b(codes)
    x = O.LUT[codes].astype(np.float64)
    hop = n // ov
    for continuous in (True, False):
        out = O.spectrometer_ex_f32(w, ntime, 1, n, ov, K, O.WIN_NONE, continuous, False)
        stride = K * hop if continuous else (K - 1) * hop + n
        nspec = ((ntime - n) // hop + 1) // K if continuous else ntime // stride
        assert out.shape[0] == nspec
        for s in range(nspec):
            segs = np.stack([x[s * stride + j * hop: s * stride + j * hop + n] for j in range(K)])
            P = np.mean(np.abs(np.fft.rfft(segs, axis=1)[:, : n // 2]) ** 2, axis=0)
            assert nrel(out[s, 0], P) < 1e-6


# ---- visibility products ---------------------------------------------------
@pytest.mark.parametrize("n", [8, 64, 512])
def test_lags_closed_forms(n):
    """Oracle lags pinned by the backward-DFT definition (brute force FP64),
    a single-delay KAT and Parseval."""
    rng = np.random.default_rng(40 + n)
    v = (rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))).astype(np.complex64)
    scale = 0.25
    got = O.lags(v, scale)
    k = np.arange(n)
    brute = np.stack([[np.sum(row.astype(np.complex128) * np.exp(2j * np.pi * k * m / n)) * scale
                       for m in range(n)] for row in v])
    assert nrel(got, brute) < 1e-5
    assert nrel(O.lags_f64(v, scale), brute) < 1e-12
    # a visibility with linear phase e^{-2 pi i k m0 / n} is a lag at m0 (a delay of m0 lags)
    m0 = n // 3
    one = np.exp(-2j * np.pi * k * m0 / n).astype(np.complex64)[None, :]
    lag = O.lags(one, 1.0 / n)[0]
    assert abs(lag[m0] - 1.0) < 1e-5 and np.max(np.abs(np.delete(lag, m0))) < 1e-5
    # Parseval: sum |lag|^2 = n scale^2 sum |v|^2
    assert abs(np.sum(np.abs(got) ** 2) / (n * scale ** 2 * np.sum(np.abs(v) ** 2)) - 1) < 1e-5


def test_amp_phase_closed_form():
    z = np.array([3 + 4j, -1 + 0j, 0 - 2j, 1e-3 + 1e-3j], np.complex64)
    a, p = O.amp_phase(z, 2.0)
    assert np.allclose(a, [10.0, 2.0, 4.0, 2 * np.sqrt(2) * 1e-3], rtol=1e-6)
    assert np.allclose(p, [np.arctan2(4, 3), np.pi, -np.pi / 2, np.pi / 4], atol=1e-6)


def _comb(ntones, f0_hz, df_hz, tau_ns, phi0=0.3, amp=1.0):
    f = f0_hz + df_hz * np.arange(ntones)
    return amp * np.exp(1j * (phi0 + 2 * np.pi * f * tau_ns * 1e-9))


@pytest.mark.parametrize("ntones,tau", [(16, 37.0), (16, -120.0), (8, 480.0), (3, 10.0), (2, -250.0)])
def test_pcal_delay_recovers_linear_phase(ntones, tau):
    """A comb whose phase grows as 2 pi f tau gives tau back (within the
    +-500/spacing[MHz] ns window; 1 MHz spacing -> +-500 ns)."""
    f0, df = 10e3, 1e6
    z = _comb(ntones, f0, df, tau)
    bc = f0 + df * (ntones - 1) / 2
    got = O.pcal_delay(z, f0, df, bc, 1e9)
    assert abs(got - tau) < 1e-9


def test_pcal_delay_wrap_band_and_degenerate():
    f0, df = 10e3, 1e6
    # beyond the window the delay wraps by the ambiguity 1000/spacing[MHz] ns
    z = _comb(16, f0, df, 700.0)
    assert abs(O.pcal_delay(z, f0, df, f0 + 7.5 * df, 1e9) - (700.0 - 1000.0)) < 1e-9
    # tones outside the valid band do not count: garbage there changes nothing
    z = _comb(16, f0, df, 42.0)
    bc, bv = f0 + 7.5 * df, 4.6e6
    ref = O.pcal_delay(z, f0, df, bc, bv)
    zz = z.copy()
    zz[:2] = np.exp(1j * np.array([2.9, -1.1]))
    zz[-2:] = np.exp(1j * np.array([0.4, 2.2]))
    assert abs(O.pcal_delay(zz, f0, df, bc, bv) - ref) < 1e-12 and abs(ref - 42.0) < 1e-9
    # fewer than two tones, or no pair inside the band
    assert O.pcal_delay(z[:1], f0, df, bc, bv) == 0.0
    assert O.pcal_delay(z, f0, df, bc, 0.1e6) == 0.0


@pytest.mark.parametrize("window", [O.WIN_NONE, O.WIN_HANN])
def test_xpol_oracle_pins(window):
    """Cross-spectrum oracle: X conj(X) of a channel with itself is the
    pinned auto spectrum (real), swapping the streams conjugates, and a copy
    delayed by d samples gives the phase slope +2 pi k d / n (the product is
    A conj(B), not conj(A) B)."""
    rng = np.random.default_rng(77)
    n, ov, K, ntime = 1024, 2, 6, 1024 * 40
    codes = rng.integers(0, 4, 2 * ntime).astype(np.uint8)
    wa = O.pack2b(codes)                                  # 2 channels interleaved
    auto = O.spectrometer_ex_f32(wa, ntime, 2, n, ov, K, window, False, True)
    x00 = O.spectrometer_xpol_f64(wa, 2, 0, wa, 2, 0, ntime, n, ov, K, window)
    assert np.max(np.abs(x00.imag)) <= 1e-12 * np.max(x00.real)
    assert np.max(np.abs(x00.real - auto[:, 0, :])) / np.max(auto[:, 0, :]) < 1e-5
    x01 = O.spectrometer_xpol_f64(wa, 2, 0, wa, 2, 1, ntime, n, ov, K, window)
    x10 = O.spectrometer_xpol_f64(wa, 2, 1, wa, 2, 0, ntime, n, ov, K, window)
    assert np.max(np.abs(x01 - np.conj(x10))) <= 1e-12 * np.max(np.abs(x01))
    # delayed copy: channel 0 of b is channel 0 of a late by d samples
    d = 3
    c0 = codes.reshape(ntime, 2)[:, 0]
    cb = np.concatenate([rng.integers(0, 4, d).astype(np.uint8), c0[:-d]])
    wb = O.pack2b(cb)
    xd = O.spectrometer_xpol_f64(wa, 2, 0, wb, 1, 0, ntime, n, ov, K, window)
    k = np.arange(20, n // 2 - 20)
    err = np.angle(xd[:, k].sum(axis=0) * np.exp(-2j * np.pi * k * d / n))
    assert np.median(np.abs(err)) < 0.05


def test_threaded_and_channelized_xengines():
    """The oracle's threaded X-engine equals the serial one bit for bit, and
    the channelized one in the generic core's loop order (input pairs outer,
    genericAddProduct_32fc_conj over the channels) gives the FP64 sums."""
    rng = np.random.default_rng(77)
    C_, T, I = 7, 33, 11
    X = (rng.standard_normal((C_, T, I)) + 1j * rng.standard_normal((C_, T, I))).astype(np.complex64)
    a = O.xengine_f32(X)
    assert np.array_equal(O.xengine_f32_mt(X, 3).view(np.uint32), a.view(np.uint32))
    ref = np.stack([(X[c].T.astype(np.complex128) @ X[c].conj().astype(np.complex128))[np.triu_indices(I)]
                    for c in range(C_)], 1)
    for nth in (1, 4):
        got = O.xcorr_tic_f32(X.transpose(1, 2, 0), nth)
        assert np.abs(got - ref).max() < 1e-5 * np.abs(ref).max()


def test_round_int_f32():
    """The integer-to-FP32 rounding of the exact 2-bit reference: numpy's
    FP64 -> FP32 conversion (round to nearest, ties to even) on integers FP64
    holds exactly, and the tie cases by hand (both neighbours, odd and even)."""
    rng = np.random.default_rng(3)
    for n in list(rng.integers(-(1 << 52), 1 << 52, 2000)) + [1, -1, (1 << 24) + 1, (1 << 25) + 3]:
        n = int(n)
        for e in (-44, -22, 0):
            assert O.round_int_f32(n, e) == float(np.float32(np.ldexp(np.float64(n), e)))
    # ties: 2^24 + 1 lies half way between 2^24 and 2^24 + 2 (even mantissa below),
    # 2^24 + 3 half way between 2^24 + 2 and 2^24 + 4 (even mantissa above)
    assert O.round_int_f32((1 << 24) + 1, 0) == float(1 << 24)
    assert O.round_int_f32((1 << 24) + 3, 0) == float((1 << 24) + 4)
    assert O.round_int_f32(-((1 << 24) + 3), 0) == -float((1 << 24) + 4)
    # beyond FP64: one more unit below the tie point rounds down, at or above up
    big = (1 << 80) + (1 << 56)            # exactly half an FP32 ulp above 2^80 (ulp 2^57)
    assert O.round_int_f32(big, -44) == 2.0 ** 36
    assert O.round_int_f32(big + 1, -44) == (2.0 ** 80 + 2.0 ** 57) * 2.0 ** -44


def _direct_exact_2bit(codes):
    """Brute force: the level values times 2^22 as Python integers, products
    summed without rounding, each sum rounded once by round_int_f32."""
    C, T, I = codes.shape
    lv = {k: int(np.ldexp(np.float64(O.LUT[k]), 22)) for k in range(4)}
    out = np.empty((I * (I + 1) // 2, C), np.complex64)
    for c in range(C):
        re = [[lv[int(codes[c, t, i]) & 3] for t in range(T)] for i in range(I)]
        im = [[lv[(int(codes[c, t, i]) >> 2) & 3] for t in range(T)] for i in range(I)]
        p = 0
        for i in range(I):
            for j in range(i, I):
                vr = sum(re[i][t] * re[j][t] + im[i][t] * im[j][t] for t in range(T))
                vi = sum(im[i][t] * re[j][t] - re[i][t] * im[j][t] for t in range(T))
                out[p, c] = O.round_int_f32(vr, -44) + 1j * O.round_int_f32(vi, -44)
                p += 1
    return out


def test_exact_2bit_reference():
    """xengine_2bit_exact_f32: one input and one spectrum gives |x|^2 = H^2 + 1
    (the FP32 level H squared exactly, then rounded once); on random codes it
    equals the integer brute force bit for bit (including 8192-spectrum sums
    whose exact values need more than FP64's 53 bits), agrees with the FP64 sums of the level
    values rounded to FP32 to within one unit in the last place, and gives
    real, positive autocorrelations."""
    Hf = np.float64(np.float32(H))
    one = np.array([[[3 | (1 << 2)]]], np.uint8)          # re = +H, im = -1
    v = O.xengine_2bit_exact_f32(one)
    assert v.shape == (1, 1) and v[0, 0] == np.complex64(np.float32(Hf * Hf + 1.0))
    codes = O.gen_channelized_codes(7, 3, 200, 9, 0.05, 0.9816)
    ex = O.xengine_2bit_exact_f32(codes)
    assert np.array_equal(ex.view(np.uint32), _direct_exact_2bit(codes).view(np.uint32))
    long_codes = O.gen_channelized_codes(9, 1, 8192, 3, 0.3, 0.9816)
    assert np.array_equal(O.xengine_2bit_exact_f32(long_codes).view(np.uint32),
                          _direct_exact_2bit(long_codes).view(np.uint32))
    X = O.LUT[codes & 3].astype(np.float64) + 1j * O.LUT[(codes >> 2) & 3].astype(np.float64)
    iu = np.triu_indices(9)
    ref = np.stack([(X[c].T @ X[c].conj())[iu] for c in range(3)], 1).astype(np.complex64)
    d = np.abs(ex.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert d.max() <= 1
    autos = [p for p, (i, j) in enumerate(zip(*iu)) if i == j]
    assert np.all(ex[autos].imag == 0) and np.all(ex[autos].real > 0)

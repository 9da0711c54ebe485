"""GPU parity of the gfx950 kernels against the CPU oracle (seeded inputs).

Bar: bit-exact for unpacked samples and generated words; normwise relative
error <= 1e-5 for FP32 spectra, visibilities and FFT outputs against the
FP32 restatement of the reference's generic-C chain; <= 1e-3 for the FP16
X-engine against the FP64 oracle.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


def nrel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def fp8_e4m3(x):
    """Round-to-nearest-even onto the OCP e4m3 grid (normal range only)."""
    x = np.asarray(x, dtype=np.float64)
    e = np.floor(np.log2(np.abs(x)))
    q = 2.0 ** (e - 3)
    return np.round(x / q) * q


@pytest.mark.parametrize("which", [0, 1, 2])
def test_mfma_layout(ctx, which):
    d = ctx.mfma_probe(which)
    v = np.arange(1, 33, dtype=np.float64)
    if which == 2:
        v = fp8_e4m3(v)
    np.testing.assert_array_equal(d, np.outer(v, v).astype(np.float32))


@pytest.mark.parametrize("nchan", [1, 2, 4, 16])
def test_unpack_bitexact(ctx, nchan):
    rng = np.random.default_rng(11 + nchan)
    ntime = 4096 + 16 // nchan * 3
    nw = (ntime * nchan + 15) // 16
    words = rng.integers(0, 2**32, nw, dtype=np.uint64).astype(np.uint32)
    dw = ctx.to_device(words)
    dout = ctx.alloc(4 * nchan * ntime)
    ctx.unpack2b(dw, ntime, nchan, dout)
    got = dout.download(np.float32, shape=(nchan, ntime))
    ref = O.unpack_1ch(words, ntime)[None, :] if nchan == 1 else O.unpack_nch(words, ntime, nchan)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_synth_station_matches_oracle(ctx):
    nsamp = 1 << 16
    dw = ctx.alloc(nsamp // 4)
    ctx.synth_station(1234, 3, 0.05, 5, 0.03, 10e3 / 16e6, 0.9816, nsamp, dw)
    got = dw.download(np.uint32)
    ref = O.gen_station(1234, 3, 0.05, 5, 0.03, 10e3 / 16e6, 0.9816, nsamp)
    assert np.mean(got == ref) > 0.9999  # FP64 libm ulp differences may flip a rare code
    assert np.sum(got != ref) <= 2


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384])
def test_fft_c2c(ctx, n):
    rng = np.random.default_rng(n)
    batch = 3
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(np.complex64)
    di = ctx.to_device(x)
    do = ctx.alloc(x.nbytes)
    ctx.fft_c2c(di, do, n, batch)
    got = do.download(np.complex64, shape=(batch, n))
    ref64 = np.fft.fft(x.astype(np.complex128), axis=1)
    assert nrel(got, ref64) < 2e-6
    ref32 = np.stack([O.fft_r2(x[b]) for b in range(batch)])
    assert nrel(got, ref32) < TOL32


@pytest.mark.parametrize("n", [4, 16, 64, 256, 1024, 8192, 16384, 32768])
@pytest.mark.parametrize("keep", [False, True])
def test_fft_r2c(ctx, n, keep):
    rng = np.random.default_rng(n + 7)
    batch = 5 if n <= 8192 else 2
    x = rng.standard_normal((batch, n)).astype(np.float32)
    di = ctx.to_device(x)
    m = n // 2 + (1 if keep else 0)
    do = ctx.alloc(batch * m * 8)
    ctx.fft_r2c(di, do, n, batch, keep)
    got = do.download(np.complex64, shape=(batch, m))
    ref64 = np.fft.rfft(x.astype(np.float64), axis=1)[:, :m]
    assert nrel(got, ref64) < 2e-6
    if n >= 4:
        ref32 = np.stack([O.rfft_r2(x[b]) for b in range(batch)])
        assert nrel(got[:, : n // 2], ref32) < TOL32


@pytest.mark.parametrize("n,nchan,k,nspec,win", [
    (1024, 16, 8, 2, O.WIN_SIN2),
    (256, 4, 5, 3, O.WIN_COSINE2),
    (4096, 16, 3, 1, O.WIN_NONE),
    (32768, 16, 4, 1, O.WIN_SIN2),
])
def test_spectrometer(ctx, n, nchan, k, nspec, win):
    ntime = n * k * nspec
    amp = np.zeros(nchan)
    freq = np.zeros(nchan)
    amp[1], freq[1] = 0.2, 37.0 / 1024
    words = O.gen_multichannel(99, nchan, amp, freq, 0.9816, ntime)
    dw = ctx.to_device(words)
    do = ctx.alloc(nspec * nchan * (n // 2) * 4)
    ctx.spectrometer(dw, ntime, nchan, n, k, win, do)
    got = do.download(np.float32, shape=(nspec, nchan, n // 2))
    ref = O.spectrometer_f32(words, ntime, nchan, n, k, win, nthreads=8)
    for s in range(nspec):
        for c in range(nchan):
            assert nrel(got[s, c], ref[s, c]) < TOL32, (s, c)


def _streams(I, nsamp, seed=7, rho=0.05):
    return [O.gen_station(seed, i, rho, i % 8, 0.03 if i == 0 else 0.0, 10e3 / 16e6, 0.9816, nsamp)
            for i in range(I)]


def _upload_streams(ctx, streams):
    w = np.stack(streams)
    return ctx.to_device(w), w.shape[1]


@pytest.mark.parametrize("n,I,nt,phase", [(1024, 4, 16, False), (8192, 6, 8, True), (256, 3, 40, True)])
def test_fstage_f32(ctx, n, I, nt, phase):
    import fxlib
    t0 = 3
    streams = _streams(I, n * (t0 + nt))
    models = [(0.37 * i + 0.1, 1e-3 * i, 0.5 * i, 0.01 * i) for i in range(I)] if phase else None
    dw, stride = _upload_streams(ctx, streams)
    corr = fxlib.Correlator(ctx, I, n, 8 * ((nt + 7) // 8), fxlib.FX_PREC_F32, use_phase=phase)
    if phase:
        corr.set_phase_model(models)
    dX = ctx.alloc((n // 2) * nt * I * 8)
    corr.fstage(dw, stride, t0, nt, dX, fxlib.FX_PREC_F32)
    got = dX.download(np.complex64, shape=(n // 2, nt, I))
    ref = O.fstage_f32(streams, n, t0, nt, models)
    assert nrel(got, ref) < TOL32
    corr.close()


@pytest.mark.parametrize("n,I,nt,phase", [(4096, 8, 16, True), (8192, 16, 8, True), (8192, 8, 16, False),
                                           (16384, 8, 8, True)])
def test_fstage_f16_direct(ctx, n, I, nt, phase):
    """The FP16-output F kernels with the corner turn folded in (one launch,
    FFT lengths 4096, 8192 and 16384): every channel of every item within 1e-3
    (normwise) of the FP32 oracle with the default FP16 power scale -- the
    FP16 budget of the visibilities -- and no channel's error beyond 1e-2 of
    that spectrum's RMS (the FP16 rounding of the output is about 3e-4)."""
    import fxlib
    t0 = 5
    streams = _streams(I, n * (t0 + nt), seed=11)
    models = [(0.37 * i + 0.1, 1e-3 * i, 0.5 * i, 0.01 * i) for i in range(I)] if phase else None
    dw, stride = _upload_streams(ctx, streams)
    corr = fxlib.Correlator(ctx, I, n, nt, fxlib.FX_PREC_F16, use_phase=phase)
    if phase:
        corr.set_phase_model(models)
    # the direct kernel (persistent: after the clear of its block counters)
    assert corr.fstage_name(nt, fxlib.FX_PREC_F16).startswith("k_fstage_direct")
    assert corr.fstage_launches(nt, fxlib.FX_PREC_F16) in (1, 2)
    dX = ctx.alloc((n // 2) * nt * I * 4)
    corr.fstage(dw, stride, t0, nt, dX, fxlib.FX_PREC_F16)
    h = dX.download(np.float16, shape=(n // 2, nt, I, 2)).astype(np.float32)
    got = h[..., 0] + 1j * h[..., 1]
    # FP16 channels carry the correlator's default power scale 1/sqrt(N E[x^2]), E[x^2] = 4.316
    ref = O.fstage_f32(streams, n, t0, nt, models) * np.float32(1.0 / np.sqrt(n * 4.316))
    err = nrel(got, ref)
    rms = np.sqrt(np.mean(np.abs(ref) ** 2, axis=0))          # per (t, i)
    worst = float(np.max(np.abs(got - ref) / rms[None]))
    print(f"FP16 F-stage n={n} I={I} nt={nt}: normwise {err:.2e}, worst channel {worst:.2e} of the spectrum RMS")
    assert err < 1e-3
    assert worst < 1e-2
    corr.close()


@pytest.mark.parametrize("I,C,T", [(4, 16, 24), (70, 8, 32), (128, 4, 16)])
def test_xengine_f32(ctx, I, C, T):
    import fxlib
    rng = np.random.default_rng(I * 1000 + C)
    X = (rng.standard_normal((C, T, I)) + 1j * rng.standard_normal((C, T, I))).astype(np.complex64)
    corr = fxlib.Correlator(ctx, I, 2 * C if 2 * C >= 256 else 256, 8, fxlib.FX_PREC_F32)
    # the X-engine only looks at the first C channels of the engine's channel
    # count; use an engine with exactly C channels when possible
    corr.close()
    corr = fxlib.Correlator(ctx, I, max(2 * C, 256), 8, fxlib.FX_PREC_F32)
    Cn = corr.nchan
    Xp = np.zeros((Cn, T, I), dtype=np.complex64)
    Xp[:C] = X
    dX = ctx.to_device(Xp)
    corr.xengine(dX, T)
    got = corr.visibilities()[:, :C]
    ref = O.xengine_f32(X)
    assert nrel(got, ref) < TOL32
    corr.close()


def test_xengine_f16_vs_f64(ctx):
    import fxlib
    I, T = 96, 64
    rng = np.random.default_rng(5)
    Cn = 128
    X = (rng.standard_normal((Cn, T, I)) + 1j * rng.standard_normal((Cn, T, I))).astype(np.complex64)
    Xh = np.empty((Cn, T, I, 2), dtype=np.float16)
    Xh[..., 0] = X.real
    Xh[..., 1] = X.imag
    corr = fxlib.Correlator(ctx, I, 2 * Cn, 8, fxlib.FX_PREC_F16)
    corr.set_scales(np.ones(Cn), np.ones(I))
    dX = ctx.to_device(Xh)
    corr.xengine(dX, T)
    got = corr.visibilities()
    ref = O.xengine_f64(X)
    assert nrel(got, ref) < 1e-3
    corr.close()


@pytest.mark.parametrize("n,I,T,chunk,phase", [(1024, 4, 64, 32, False), (512, 10, 40, 16, True)])
def test_corr_end_to_end_f32(ctx, n, I, T, chunk, phase):
    import fxlib
    streams = _streams(I, n * T)
    models = [(1.5 * i, 2e-4 * i, 0.3 * i, 0.02) for i in range(I)] if phase else None
    dw, stride = _upload_streams(ctx, streams)
    corr = fxlib.Correlator(ctx, I, n, chunk, fxlib.FX_PREC_F32, use_phase=phase)
    if phase:
        corr.set_phase_model(models)
    corr.process(dw, stride, 0, T)
    got = corr.visibilities()
    ref = O.corr_f32(streams, n, T, models, nthreads=4)
    assert nrel(got, ref) < TOL32
    ref64 = O.corr_f64(streams, n, T, models, nthreads=4)
    assert nrel(got, ref64) < TOL32
    corr.close()


def test_corr_f16_vs_f64(ctx):
    import fxlib
    n, I, T = 1024, 12, 64
    streams = _streams(I, n * T, rho=0.2)
    dw, stride = _upload_streams(ctx, streams)
    corr = fxlib.Correlator(ctx, I, n, 32, fxlib.FX_PREC_F16)
    corr.process(dw, stride, 0, T)
    got = corr.visibilities()
    ref64 = O.corr_f64(streams, n, T, None, nthreads=4)
    assert nrel(got, ref64) < 1e-3
    corr.close()


@pytest.mark.parametrize("bw,spacing,offset", [(8e6, 1e6, 10e3), (8e6, 1e6, 0), (8e6, 1e6, 7e3)])
def test_pcal_2bit(ctx, bw, spacing, offset):
    import fxlib
    fs = 2 * bw
    nsamp = 1 << 18
    words = O.gen_station(77, 0, 0.0, 0, 0.5, (offset + 2 * spacing) / fs, 0.9816, nsamp)
    dw = ctx.to_device(words)
    pc = fxlib.PCal(ctx, bw, spacing, offset)
    cfg = O.pcal_config(bw, spacing, offset)
    assert (pc.kind, pc.nbins, pc.ntones) == (cfg.kind, cfg.nbins, cfg.ntones)
    half = nsamp // 2
    pc.extract_2bit(dw, half)
    dw2 = ctx.to_device(words[half // 16:])
    pc.extract_2bit(dw2, nsamp - half)
    got, n = pc.final()
    assert n == nsamp
    x = O.unpack_1ch(words, nsamp)
    acc = O.pcal_extract(cfg, x, 0)
    ref = O.pcal_final(cfg, acc)
    assert nrel(got, ref) < 1e-12
    pc.close()


@pytest.mark.parametrize("bw,spacing,offset", [(8e6, 1e6, 10e3), (8e6, 1e6, 0)])
def test_pcal_2bit_inputs(ctx, bw, spacing, offset):
    """One multi-input extractor (one fold + one final launch for all inputs)
    against the oracle per input, over two extract calls."""
    import fxlib
    fs = 2 * bw
    nsamp, ninp = 1 << 17, 3
    rows = [O.gen_station(90 + i, 0, 0.0, 0, 0.5, (offset + (i + 1) * spacing) / fs, 0.9816, nsamp)
            for i in range(ninp)]
    nw = nsamp // 16
    stride = nw + 64
    buf = np.zeros((ninp, stride), dtype=np.uint32)
    for i in range(ninp):
        buf[i, :nw] = rows[i]
    half = nsamp // 2
    pc = fxlib.PCal(ctx, bw, spacing, offset, ninp=ninp)
    pc.extract_2bit_inputs(ctx.to_device(buf.ravel()), stride, half)
    tail = np.ascontiguousarray(buf[:, half // 16:])
    pc.extract_2bit_inputs(ctx.to_device(tail.ravel()), tail.shape[1], nsamp - half)
    got, n = pc.final()
    assert n == nsamp and got.shape == (ninp, pc.ntones)
    cfg = O.pcal_config(bw, spacing, offset)
    for i in range(ninp):
        ref = O.pcal_final(cfg, O.pcal_extract(cfg, O.unpack_1ch(rows[i], nsamp), 0))
        assert nrel(got[i], ref) < 1e-12, i
    pc.close()


def test_memset_fill_and_bytes(ctx):
    """fx_memset: aligned whole words go through the fill kernel, other
    ranges through the byte memset; both give the byte pattern and touch
    nothing outside the range."""
    import ctypes as C
    import fxlib
    lib = fxlib.load()
    buf = ctx.alloc(4096)
    buf.upload(np.full(4096, 0x11, np.uint8))
    base = buf.ptr.value
    assert lib.fx_memset(ctx.handle, C.c_void_p(base + 1024), 0x5A, 2048) == 0      # fill kernel
    assert lib.fx_memset(ctx.handle, C.c_void_p(base + 3), 0xC3, 7) == 0            # byte memset
    assert lib.fx_memset(ctx.handle, C.c_void_p(base + 4000), 0x1FF, 8) == 0        # low byte only
    got = buf.download(np.uint8)
    want = np.full(4096, 0x11, np.uint8)
    want[1024:3072] = 0x5A
    want[3:10] = 0xC3
    want[4000:4008] = 0xFF
    assert np.array_equal(got, want)
    buf.free()

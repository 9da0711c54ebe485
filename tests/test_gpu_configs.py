"""Parity at the benchmark configurations' real shapes.

* mini FX correlator: 4 stations, 2^20 samples each, correlated fraction 0.05
  with seeded integer delays 0-7, a 10 kHz phase-cal tone at 0.03 sigma
  (fs = 16 MHz), 1024-point FFT, all 10 products over all 1024 spectra --
  FP32 engine and the minicorrelator control-block drop-in against the FP32
  oracle (<= 1e-5) and the FP64 oracle; phase-cal tones in FP64 against the
  oracle;
* spectrometer: one station, 16 interleaved subbands, 2^28 samples, 32768-point
  FFT, power over the whole stream, 0.2 sigma tones in 3 seeded subbands --
  against the FP32 oracle (<= 1e-5) and the tone peaks at the seeded bins;
* mid-size correlator shape: 128 inputs, 8192-point FFT, FP64 phase model
  with fractional delays 0-50 samples, 16 spectra -- FP32 against the FP32
  oracle, FP16 against the FP64 oracle (<= 1e-3); and the whole 65536-spectrum
  integration: FP16 against FP32, autos, split-run additivity, and the last
  64 spectra against the oracle;
* large-N shape: 512 inputs, 8192 spectra of 2-bit channelized voltages
  (128 channels here), exact FP8 engine and FP16 engine against FP64 sums on
  sampled channels (the exact FP8 and FP4 engines bit for bit against the
  oracle's exact integer sums rounded once); and all 1024 channels of the large-N workload
  on a 64-spectrum time slice through the exact FP8 and FP4, FP16 and plain
  FP32 engines, every product and channel against FP64 sums and the exact
  engines bit for bit against the exact sums (FP32 also against the FP32
  restatement on sampled channels).
"""
import numpy as np
import pytest

import fxlib
import oracle as O
import tasklib as TL

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
FS = 16e6


def nrel(a, b):
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_mini_correlator_full(ctx):
    nst, nsamp, n = 4, 1 << 20, 1024
    T = nsamp // n
    delays = np.random.default_rng(2024).integers(0, 8, nst)
    streams = [O.gen_station(2024, s, 0.05, int(delays[s]), 0.03, 10e3 / FS, 0.9816, nsamp)
               for s in range(nst)]
    words = np.stack(streams)
    dw = ctx.to_device(words)
    corr = fxlib.Correlator(ctx, nst, n, 256, fxlib.FX_PREC_F32)
    corr.process(dw, words.shape[1], 0, T)
    got = corr.visibilities()
    corr.close()
    ref32 = O.corr_f32(streams, n, T, nthreads=8)
    ref64 = O.corr_f64(streams, n, T, nthreads=8)
    assert nrel(got, ref32) < TOL32
    print(f"mini correlator: vs FP32 oracle {nrel(got, ref32):.2e}, vs FP64 {nrel(got, ref64):.2e}")
    assert nrel(got, ref64) < TOL32
    # the same through the minicorrelator control blocks, one dump
    autos, bls, st = TL.minicorr_run(streams, T)
    assert st.dumps == 1
    p = 0
    for i in range(nst):
        for j in range(i, nst):
            buf = autos[i][0] if i == j else bls[i * nst - i * (i + 1) // 2 + (j - i - 1)][0]
            assert nrel(buf[0, : n // 2] + 1j * buf[1, : n // 2], ref32[p]) < TOL32
            p += 1
    # phase-cal of the 10 kHz tone (1 MHz comb spacing, 8 MHz band), FP64
    for s in range(nst):
        pc = fxlib.PCal(ctx, FS / 2, 1e6, 10e3)
        pc.extract_2bit(ctx.to_device(streams[s]), nsamp)
        tones, used = pc.final()
        cfg = O.pcal_config(FS / 2, 1e6, 10e3)
        ref = O.pcal_final(cfg, O.pcal_extract(cfg, O.unpack_1ch(streams[s], nsamp)))
        assert used == nsamp and nrel(tones, ref) < 1e-12
        # the injected tone stands far above the other tone bins of noise
        assert abs(tones[0]) > 5 * np.median(np.abs(tones[1:]))


def test_spectrometer_full():
    nch, ntotal, n = 16, 1 << 28, 32768
    ntime = ntotal // nch
    K = ntime // n
    rng = np.random.default_rng(11)
    subs = rng.choice(nch, 3, replace=False)
    bins = rng.integers(200, n // 2 - 200, 3)
    amp = np.zeros(nch)
    freq = np.zeros(nch)
    amp[subs] = 0.2
    freq[subs] = bins / n
    with fxlib.Context(0) as c:
        # device generator (bit-identical to the oracle's, checked elsewhere)
        dw = c.alloc(ntime * nch // 4)
        c.synth_multichannel(31, nch, amp, freq, 0.9816, ntime, dw)
        words = dw.download(np.uint32)
        assert np.array_equal(words[:4096], O.gen_multichannel(31, nch, amp, freq, 0.9816, 4096 * 16 // nch))
        do = c.alloc(nch * (n // 2) * 4)
        c.spectrometer(dw, ntime, nch, n, K, fxlib.FX_WIN_SIN2, do)
        got = do.download(np.float32, shape=(1, nch, n // 2))
    ref = O.spectrometer_f32(words, ntime, nch, n, K, O.WIN_SIN2, nthreads=16)
    err = max(nrel(got[0, ch], ref[0, ch]) for ch in range(nch))
    print(f"spectrometer 16 x 16384: worst subband error vs FP32 oracle {err:.2e}")
    assert err < TOL32
    for sb, b in zip(subs, bins):
        assert int(np.argmax(got[0, sb, 1:])) + 1 == b


def test_mid_correlator_shape(ctx):
    stations, n, T = 64, 8192, 16
    I = 2 * stations
    rng = np.random.default_rng(20240229)
    tau = rng.uniform(0.0, 50.0, stations)
    models = []
    for s in range(stations):
        for _ in range(2):
            models.append((float(tau[s]), 1e-7 * s, float(rng.uniform(-np.pi, np.pi)), 1e-4))
    streams = [O.gen_station(77, i, 0.02, int(tau[i // 2]), 0.0, 0.0, 0.9816, n * T) for i in range(I)]
    words = np.stack(streams)
    dw = ctx.to_device(words)
    out = {}
    for prec in (fxlib.FX_PREC_F32, fxlib.FX_PREC_F16):
        corr = fxlib.Correlator(ctx, I, n, 16, prec, use_phase=True)
        corr.set_phase_model(models)
        corr.process(dw, words.shape[1], 0, T)
        out[prec] = corr.visibilities()
        corr.close()
    ref32 = O.corr_f32(streams, n, T, models, nthreads=16)
    ref64 = O.corr_f64(streams, n, T, models, nthreads=16)
    e32, e16 = nrel(out[0], ref32), nrel(out[1], ref64)
    print(f"mid shape: f32 vs FP32 oracle {e32:.2e}, f16 vs FP64 {e16:.2e}")
    assert e32 < TOL32
    assert e16 < 1e-3


def test_mid_correlator_full_run(ctx):
    """The mid-size correlator over its whole integration, as the benchmark
    runs it: 128 inputs x 2^29 samples (16 GiB of 2-bit words generated on the
    device), 8192-point FFT, 65536 spectra in 16384-spectrum chunks, seeded
    delays 0-50 samples with delay and fringe rates.  Size-independent checks
    at full size: the FP16 chain against the FP32 chain over all 8256 products
    x 4096 channels (<= 1e-3); autos real and positive; the run split at
    spectrum 32768 into two runs sums to the whole run; and a 64-spectrum
    window at the end of the integration (t = 65472 .., where the FP64 phase
    model's rate terms are largest) against the oracle on the same words, with
    the model's constant terms moved to the window's start (FP32 <= 1e-5
    against the FP32 oracle, FP16 <= 1e-3 against FP64)."""
    stations, n, T, chunk = 64, 8192, 65536, 16384
    I, C = 2 * stations, n // 2
    P = I * (I + 1) // 2
    rng = np.random.default_rng(20240229)
    tau = rng.uniform(0.0, 50.0, stations)
    fr0 = rng.uniform(-np.pi, np.pi, stations)
    fr1 = rng.uniform(-1e-3, 1e-3, stations)
    models = [(float(tau[s]), 1e-7 * s, float(fr0[s]), float(fr1[s])) for s in range(stations) for _ in range(2)]
    nw = T * n // 16
    dw = ctx.alloc(I * nw * 4)
    lib = fxlib.load()
    for i in range(I):
        fxlib.check(lib.fx_synth_station(ctx.handle, 20240229, i, 0.02, int(tau[i // 2]), 0.0, 0.0, 0.9816, T * n,
                                         dw.ptr.value + i * nw * 4))
    ctx.sync()

    def run(prec, t0, nt, ch):
        corr = fxlib.Correlator(ctx, I, n, ch, prec, use_phase=True)
        corr.set_phase_model(models)
        corr.process(dw, nw, t0, nt)
        v = corr.visibilities()
        corr.close()
        return v

    v16 = run(fxlib.FX_PREC_F16, 0, T, chunk)
    v32 = run(fxlib.FX_PREC_F32, 0, T, chunk)
    e = nrel(v16, v32)
    ia = np.arange(I)
    autos = ia * I - ia * (ia - 1) // 2
    a16, a32 = v16[autos].astype(np.complex128), v32[autos].astype(np.complex128)
    va = run(fxlib.FX_PREC_F16, 0, T // 2, chunk)
    vb = run(fxlib.FX_PREC_F16, T // 2, T // 2, chunk)
    esplit = nrel(va.astype(np.complex128) + vb, v16)
    del va, vb
    print(f"mid full run: f16 vs f32 {e:.2e}, split vs whole {esplit:.2e}")
    assert e < 1e-3
    assert esplit < 1e-6
    for a in (a16, a32):
        assert np.all(a.real > 0)
        assert np.max(np.abs(a.imag) / a.real) < 1e-5
    del v16, v32
    # the window at the end of the integration
    tw, nt = T - 64, 64
    wpt = n // 16
    streams = [dw.download_at((i * nw + tw * wpt) * 4, np.uint32, nt * wpt) for i in range(I)]
    shifted = [(d0 + d1 * tw, d1, f0 + f1 * tw, f1) for (d0, d1, f0, f1) in models]
    g32 = run(fxlib.FX_PREC_F32, tw, nt, nt)
    g16 = run(fxlib.FX_PREC_F16, tw, nt, nt)
    dw.free()
    r32 = O.corr_f32(streams, n, nt, shifted, nthreads=16)
    r64 = O.corr_f64(streams, n, nt, shifted, nthreads=16)
    e32, e16 = nrel(g32, r32), nrel(g16, r64)
    print(f"mid full run, last 64 spectra: f32 vs FP32 oracle {e32:.2e}, f16 vs FP64 {e16:.2e}")
    assert e32 < TOL32
    assert e16 < 1e-3


def test_large_n_shape(ctx):
    I, C, T = 512, 128, 8192
    codes = ctx.alloc(C * I * T)
    ctx.synth_channelized(5, C, T, I, 0.01, 0.9816, codes, input_major=True)
    c8 = fxlib.Correlator(ctx, I, 2 * C, T, fxlib.FX_PREC_FP8_2BIT)
    c8.xengine(codes, T)
    v8 = c8.visibilities()
    c8.close()
    c4 = fxlib.Correlator(ctx, I, 2 * C, T, fxlib.FX_PREC_FP4_2BIT)
    c4.xengine(codes, T)
    v4 = c4.visibilities()
    c4.close()
    x16 = ctx.alloc(C * T * I * 4)
    ctx.codes_to_chan(codes, True, C, T, I, fxlib.FX_PREC_F16, x16)
    c16 = fxlib.Correlator(ctx, I, 2 * C, 8, fxlib.FX_PREC_F16)
    c16.set_scales(np.ones(C, np.float32), np.ones(I, np.float32))
    c16.xengine(x16, T)
    v16 = c16.visibilities()
    c16.close()
    lut = O.LUT.astype(np.float64)
    iu = np.triu_indices(I)
    for ch in (0, 77):
        cb = codes.download_at(ch * I * T, np.uint8, I * T).reshape(I, T)
        x = lut[cb & 3] + 1j * lut[(cb >> 2) & 3]
        ref = (x @ x.conj().T)[iu]
        e8, e16 = nrel(v8[:, ch], ref), nrel(v16[:, ch], ref)
        print(f"large-N channel {ch}: fp8 exact {e8:.2e}, f16 {e16:.2e}")
        assert e8 < 1e-6
        # 8192-spectrum sums: the exact integer sums rounded once, bit for bit
        exact = O.xengine_2bit_exact_f32(cb.T[None])[:, 0]
        assert np.array_equal(np.ascontiguousarray(v8[:, ch]).view(np.uint32), exact.view(np.uint32))
        assert np.array_equal(np.ascontiguousarray(v4[:, ch]).view(np.uint32), exact.view(np.uint32))
        assert e16 < 1e-3


def test_large_n_all_channels_time_slice(ctx):
    I, C, T = 512, 1024, 64
    P = I * (I + 1) // 2
    codes = ctx.alloc(C * I * T)
    ctx.synth_channelized(2025, C, T, I, 0.01, 0.9816, codes, input_major=True)
    cb = codes.download(np.uint8).reshape(C, I, T)
    lut = O.LUT.astype(np.float64)
    iu = np.triu_indices(I)
    ref = np.empty((P, C), np.complex128)
    for c in range(C):
        x = lut[cb[c] & 3] + 1j * lut[(cb[c] >> 2) & 3]
        ref[:, c] = (x @ x.conj().T)[iu]
    vis = ctx.alloc(P * C * 8)
    got = {}
    for name, prec in (("fp8_2bit", fxlib.FX_PREC_FP8_2BIT), ("fp4_2bit", fxlib.FX_PREC_FP4_2BIT)):
        c8 = fxlib.Correlator(ctx, I, 2 * C, T, prec)
        c8.xengine(codes, T)
        c8.dump(vis)
        c8.close()
        got[name] = vis.download(np.complex64, shape=(P, C))
    for name, prec, esz in (("f16", fxlib.FX_PREC_F16, 4), ("f32", fxlib.FX_PREC_F32, 8)):
        dx = ctx.alloc(C * T * I * esz)
        ctx.codes_to_chan(codes, True, C, T, I, prec, dx)
        cc = fxlib.Correlator(ctx, I, 2 * C, 8, prec)
        cc.set_scales(np.ones(C, np.float32), np.ones(I, np.float32))
        assert cc.xengine_name(T, dx) == ("k_xengine_ring2<F16>" if name == "f16" else "k_xengine_ring2<F32>")
        cc.xengine(dx, T)
        cc.dump(vis)
        cc.close()
        dx.free()
        got[name] = vis.download(np.complex64, shape=(P, C))
    codes.free()
    vis.free()
    err = {k: nrel(v, ref) for k, v in got.items()}
    print("large-N, all 1024 channels x 64 spectra vs FP64: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()))
    assert err["fp8_2bit"] < 1e-6
    # the exact engines: the exact sums rounded once to FP32, bit for bit, on every channel
    exact = O.xengine_2bit_exact_f32(cb.transpose(0, 2, 1))
    assert np.array_equal(got["fp8_2bit"].view(np.uint32), exact.view(np.uint32))
    assert np.array_equal(got["fp4_2bit"].view(np.uint32), exact.view(np.uint32))
    assert err["f32"] < TOL32
    assert err["f16"] < 1e-3
    # the FP32 engine against the FP32 restatement of the reference MAC
    for c in (0, 511, 1023):
        x = (lut[cb[c] & 3] + 1j * lut[(cb[c] >> 2) & 3]).T.astype(np.complex64)   # [t][i]
        r32 = O.xengine_f32(x[None])[:, 0]
        assert nrel(got["f32"][:, c], r32) < TOL32

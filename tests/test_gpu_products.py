"""GPU parity of the visibility products downstream of the dump: lag spectra
(zerocorr.c:391,633-645), amplitude/phase (zerocorr.c:628-643) and the
phase-cal delay fit (m5pcal.c:94-150), against the oracle.

Bar: lags <= 1e-5 normwise vs the FP32 radix-2 restatement; amplitude and
phase to FP32 rounding of FP64 values; delays to 1e-9 ns of the restatement.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def nrel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.mark.parametrize("nchan,nrows", [(64, 37), (1024, 9), (4096, 5), (16384, 2)])
def test_lags_vs_oracle(ctx, nchan, nrows):
    rng = np.random.default_rng(nchan + nrows)
    v = (rng.standard_normal((nrows, nchan)) + 1j * rng.standard_normal((nrows, nchan))).astype(np.complex64)
    dv = ctx.to_device(v.ravel())
    dl = ctx.alloc(v.nbytes)
    scale = 1.0 / 640.0
    ctx.lags(dv, nrows, nchan, scale, dl)
    got = dl.download(np.complex64, nrows * nchan).reshape(nrows, nchan)
    assert nrel(got, O.lags(v, scale)) < 1e-5
    assert nrel(got, O.lags_f64(v, scale)) < 1e-5


def test_lags_empty_and_bad_sizes(ctx):
    import fxlib
    d = ctx.alloc(8 * 64)
    ctx.lags(d, 0, 64, 1.0, d)          # no rows: nothing to do
    with pytest.raises(fxlib.FxError):
        ctx.lags(d, 1, 48, 1.0, d)      # not a power of two
    with pytest.raises(fxlib.FxError):
        ctx.lags(d, 1, 32768, 1.0, d)   # above 16384


def test_amp_phase_vs_oracle(ctx):
    rng = np.random.default_rng(5)
    z = (rng.standard_normal(100003) + 1j * rng.standard_normal(100003)).astype(np.complex64)
    z[:3] = [0, -1, 1j]
    dz = ctx.to_device(z)
    da, dp = ctx.alloc(4 * z.size), ctx.alloc(4 * z.size)
    ctx.amp_phase(dz, z.size, 0.5, da, dp)
    a = da.download(np.float32, z.size)
    p = dp.download(np.float32, z.size)
    ra, rp = O.amp_phase(z, 0.5)
    np.testing.assert_allclose(a, ra.astype(np.float32), rtol=1e-7, atol=0)
    np.testing.assert_allclose(p, rp.astype(np.float32), rtol=1e-7, atol=1e-7)
    # one output only
    ctx.amp_phase(dz, z.size, 0.5, None, dp)
    np.testing.assert_array_equal(dp.download(np.float32, z.size), p)


def test_pcal_delay_vs_oracle(ctx):
    rng = np.random.default_rng(8)
    f0, df = 10e3, 1e6
    cases = [(16, 1e9), (16, 4.6e6), (8, 1e9), (3, 1e9), (2, 1e9), (1, 1e9)]
    for ntones, bv in cases:
        ninp = 7
        taus = rng.uniform(-900, 900, ninp)
        tones = np.stack([np.exp(1j * (rng.uniform(-3, 3) + 2 * np.pi * (f0 + df * np.arange(ntones)) * t * 1e-9
                                        + 0.2 * rng.standard_normal(ntones)))
                          * rng.uniform(0.5, 2.0) for t in taus])
        bc = f0 + df * (ntones - 1) / 2
        dt = ctx.to_device(tones.astype(np.complex128).ravel())
        dd = ctx.alloc(8 * ninp)
        ctx.pcal_delay(dt, ninp, ntones, f0, df, bc, bv, dd)
        got = dd.download(np.float64, ninp)
        ref = np.array([O.pcal_delay(tones[y], f0, df, bc, bv) for y in range(ninp)])
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-9)


def test_pcal_delay_from_extracted_tones(ctx):
    """Tones from the device phase-cal extractor of a multi-tone comb with a
    known group delay: the fit returns the delay."""
    import fxlib
    bw, spacing, offset = 8e6, 1e6, 10e3
    fs = 2 * bw
    n = 1 << 18
    tau_ns = 63.0
    t = np.arange(n) / fs
    f = offset + spacing * np.arange(8)
    x = np.zeros(n)
    for fk in f:
        x += np.cos(2 * np.pi * fk * (t - tau_ns * 1e-9))
    rng = np.random.default_rng(3)
    x = x * 0.3 + rng.standard_normal(n)
    codes = np.digitize(x, [-0.9816, 0.0, 0.9816]).astype(np.uint8)
    words = O.pack2b(codes)
    pc = fxlib.PCal(ctx, bw, spacing, offset)
    pc.extract_2bit(ctx.to_device(words), n)
    tones, _ = pc.final()
    pc.close()
    dt = ctx.to_device(tones.astype(np.complex128))
    dd = ctx.alloc(8)
    ctx.pcal_delay(dt, 1, tones.size, offset, spacing, offset + spacing * 3.5, 1e9, dd)
    got = float(dd.download(np.float64, 1)[0])
    assert abs(got - O.pcal_delay(tones, offset, spacing, offset + spacing * 3.5, 1e9)) < 1e-9
    # the extractor correlates with e^{-i 2 pi f t} (PCal.cpp:139-175), so a
    # signal delayed by tau has tone phases -2 pi f tau: the fit gives -tau
    assert abs(got + tau_ns) < 3.0


def test_lags_of_correlated_pair(ctx):
    """End to end: the cross-visibility of a station and its delayed copy has
    its lag peak at the delay (in lag units of 2 samples over C channels)."""
    import fxlib
    n, nt, d = 1024, 64, 6
    C = n // 2
    a = O.gen_station(21, 0, 0.9, 0, 0.0, 0.0, 0.9816, n * nt + 64)
    b = O.gen_station(21, 1, 0.9, d, 0.0, 0.0, 0.9816, n * nt + 64)
    dw = ctx.to_device(np.stack([a, b]))
    corr = fxlib.Correlator(ctx, 2, n, nt, fxlib.FX_PREC_F32)
    corr.process(dw, a.size, 0, nt)
    vis = corr.visibilities()                    # [3][C]: (0,0), (0,1), (1,1)
    corr.close()
    dv = ctx.to_device(vis[1].astype(np.complex64))
    dl = ctx.alloc(8 * C)
    ctx.lags(dv, 1, C, 1.0 / nt, dl)
    lag = np.abs(dl.download(np.complex64, C))
    peak = int(np.argmax(lag))
    assert peak in (d // 2, C - d // 2)
    assert lag[peak] > 10 * np.median(lag)

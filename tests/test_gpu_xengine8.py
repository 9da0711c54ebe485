"""GPU parity of the 8-bit X-engines (v_mfma_scale_f32_32x32x64_f8f6f4).

* FX_PREC_FP8_2BIT / FX_PREC_FP4_2BIT: 2-bit channelized voltages through the
  exact decomposition of the level products into three integer sums on e4m3
  or e2m1 operands; the dump combines the sums exactly and rounds once, so the
  visibilities equal the oracle's exact sums rounded to FP32
  (xengine_2bit_exact_f32) bit for bit, the two operand forms are
  bit-identical, and both are within one ulp of the FP64 brute force.
* FX_PREC_FP8: e4m3 inputs; the engine must reproduce the FP64 sums of the
  decoded e4m3 values (products are exact, FP32 accumulation, <= 1e-6), and
  the full F->X chain in FP8 is reported against the FP64 oracle.
"""
import numpy as np
import pytest

import fxlib
import oracle as O

pytestmark = pytest.mark.gpu


def nrel(a, b):
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def e4m3_decode(b):
    b = np.asarray(b, np.uint8).astype(np.int64)
    s = np.where(b >> 7, -1.0, 1.0)
    e = (b >> 3) & 15
    m = b & 7
    v = np.where(e == 0, m / 8.0 * 2.0**-6, (1 + m / 8.0) * 2.0 ** (e - 7))
    return s * v


def codes_to_values(codes):
    return O.LUT[codes & 3].astype(np.float64) + 1j * O.LUT[(codes >> 2) & 3].astype(np.float64)


@pytest.mark.parametrize("nchan,nt,ninp", [(128, 64, 8)])
def test_synth_channelized_layouts(ctx, nchan, nt, ninp):
    ref = O.gen_channelized_codes(99, nchan, nt, ninp, 0.2, 0.9816)      # [c][t][i]
    d = ctx.alloc(nchan * nt * ninp)
    ctx.synth_channelized(99, nchan, nt, ninp, 0.2, 0.9816, d, input_major=False)
    assert np.array_equal(d.download(np.uint8).reshape(nchan, nt, ninp), ref)
    ctx.synth_channelized(99, nchan, nt, ninp, 0.2, 0.9816, d, input_major=True)
    im = d.download(np.uint8).reshape(nchan, ninp, nt)
    assert np.array_equal(im, ref.transpose(0, 2, 1))
    out = ctx.alloc(nchan * nt * ninp * 8)
    ctx.codes_to_chan(d, True, nchan, nt, ninp, fxlib.FX_PREC_F32, out)
    got = out.download(np.complex64).reshape(nchan, nt, ninp)
    assert np.array_equal(got, codes_to_values(ref).astype(np.complex64))


@pytest.mark.parametrize("ninp,nt", [(40, 80), (200, 64), (256, 48)])
def test_2bit_exact(ctx, ninp, nt):
    nchan = 128
    codes = O.gen_channelized_codes(7 + ninp, nchan, nt, ninp, 0.05, 0.9816)   # [c][t][i]
    d = ctx.to_device(np.ascontiguousarray(codes.transpose(0, 2, 1)))
    got = {}
    for prec in (fxlib.FX_PREC_FP8_2BIT, fxlib.FX_PREC_FP4_2BIT):
        corr = fxlib.Correlator(ctx, ninp, 2 * nchan, 64, prec)
        corr.xengine(d, nt)
        got[prec] = corr.visibilities()
        corr.close()
    ref = O.xengine_f64(codes_to_values(codes).astype(np.complex64))
    g8, g4 = got[fxlib.FX_PREC_FP8_2BIT], got[fxlib.FX_PREC_FP4_2BIT]
    assert np.array_equal(g8.view(np.uint32), g4.view(np.uint32))
    exact = O.xengine_2bit_exact_f32(codes)
    assert np.array_equal(g4.view(np.uint32), exact.view(np.uint32))
    r32 = ref.astype(np.complex64)
    ulp = np.abs(g4.view(np.int32).astype(np.int64) - r32.view(np.int32).astype(np.int64))
    print(f"2-bit ninp={ninp} nt={nt}: normwise error vs FP64 {nrel(g4, ref):.2e}, "
          f"{100 * np.mean(ulp == 0):.4f}% of components equal to the FP64 brute force rounded to FP32")
    assert ulp.max() <= 1


@pytest.mark.parametrize("prec", [fxlib.FX_PREC_FP8_2BIT, fxlib.FX_PREC_FP4_2BIT])
def test_2bit_accumulates_over_calls(ctx, prec):
    nchan, ninp, nt = 128, 64, 32
    codes = O.gen_channelized_codes(3, nchan, 2 * nt, ninp, 0.1, 0.9816)
    corr = fxlib.Correlator(ctx, ninp, 2 * nchan, 64, prec)
    for h in range(2):
        part = np.ascontiguousarray(codes[:, h * nt:(h + 1) * nt, :].transpose(0, 2, 1))
        corr.xengine(ctx.to_device(part), nt)
    got = corr.visibilities()
    corr.reset()
    assert np.all(corr.visibilities() == 0)
    corr.close()
    assert np.array_equal(got.view(np.uint32), O.xengine_2bit_exact_f32(codes).view(np.uint32))


@pytest.mark.parametrize("ninp,nt", [(48, 64), (160, 40)])
def test_fp8_engine_sums(ctx, ninp, nt):
    nchan = 128
    rng = np.random.default_rng(ninp)
    # random finite e4m3 bytes of moderate magnitude (|x| in [2^-3, 8])
    e = rng.integers(4, 10, (nchan, ninp, nt, 2))
    m = rng.integers(0, 8, (nchan, ninp, nt, 2))
    s = rng.integers(0, 2, (nchan, ninp, nt, 2))
    b = ((s << 7) | (e << 3) | m).astype(np.uint8)                 # [c][i][t][re,im]
    d = ctx.to_device(b)
    corr = fxlib.Correlator(ctx, ninp, 2 * nchan, 64, fxlib.FX_PREC_FP8)
    corr.set_scales(np.ones(nchan, np.float32), np.ones(ninp, np.float32))
    corr.xengine(d, nt)
    got = corr.visibilities()
    corr.close()
    v = e4m3_decode(b[..., 0]) + 1j * e4m3_decode(b[..., 1])         # [c][i][t]
    X = np.ascontiguousarray(v.transpose(0, 2, 1))                    # [c][t][i]
    ref = np.zeros((ninp * (ninp + 1) // 2, nchan), np.complex128)
    iu = np.triu_indices(ninp)
    for c in range(nchan):
        V = X[c].T @ np.conj(X[c])                                    # V[i][j] = sum_t X_i conj(X_j)
        ref[:, c] = V[iu]
    # the e4m3 products are exact, but the MFMA's 64-term dot-product adder
    # keeps fewer bits than FP32 when the terms span a wide exponent range
    # (here 2^-6 .. 2^8): measured 1.5e-5, a small truncation bias
    err = nrel(got, ref)
    print(f"FP8 engine vs FP64 sums of the e4m3 values: {err:.2e}")
    assert err < 1e-4


def test_fp8_fx_chain_vs_fp64(ctx):
    n, ninp, nspec = 1024, 24, 64
    streams = [O.gen_station(41, i, 0.3, i % 8, 0.0, 0.0, 0.9816, n * nspec) for i in range(ninp)]
    words = np.stack(streams)
    dw = ctx.to_device(words)
    ref = O.corr_f64(streams, n, nspec, nthreads=4)
    errs = {}
    for prec in (fxlib.FX_PREC_F32, fxlib.FX_PREC_F16, fxlib.FX_PREC_FP8):
        corr = fxlib.Correlator(ctx, ninp, n, 32, prec)
        corr.process(dw, words.shape[1], 0, nspec)
        errs[prec] = nrel(corr.visibilities(), ref)
        corr.close()
    print(f"F->X vs FP64 ({nspec} spectra): f32 {errs[0]:.2e}  f16 {errs[1]:.2e}  fp8 {errs[2]:.2e}")
    assert errs[fxlib.FX_PREC_F32] < 1e-5
    assert errs[fxlib.FX_PREC_F16] < 1e-3
    # e4m3 keeps 3 mantissa bits: about 2^-5 relative rounding per sample
    assert errs[fxlib.FX_PREC_FP8] < 5e-2


def test_fp8_fstage_layout(ctx):
    # the FP8 F-stage writes [c][i][t] e4m3 pairs of the scaled FP32 F-stage
    n, ninp, nspec = 512, 8, 16
    streams = [O.gen_station(5, i, 0.1, 0, 0.0, 0.0, 0.9816, n * nspec) for i in range(ninp)]
    words = np.stack(streams)
    dw = ctx.to_device(words)
    corr = fxlib.Correlator(ctx, ninp, n, 16, fxlib.FX_PREC_FP8)
    corr.set_scales(np.full(n // 2, 0.25, np.float32), np.ones(ninp, np.float32))
    out = ctx.alloc((n // 2) * ninp * nspec * 2)
    corr.fstage(dw, words.shape[1], 0, nspec, out, fxlib.FX_PREC_FP8)
    b = out.download(np.uint8).reshape(n // 2, ninp, nspec, 2)
    corr.close()
    got = e4m3_decode(b[..., 0]) + 1j * e4m3_decode(b[..., 1])
    X = O.fstage_f32(streams, n, 0, nspec).astype(np.complex128) * 0.25     # [c][t][i]
    want = X.transpose(0, 2, 1)
    # round-to-nearest e4m3: relative error <= 2^-4 per component (normal range)
    for g, w in ((got.real, want.real), (got.imag, want.imag)):
        mag = np.maximum(np.abs(w), 2.0**-6)
        assert np.all(np.abs(g - w) <= mag * 2.0**-4 + 2.0**-10)

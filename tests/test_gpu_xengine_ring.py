"""GPU parity of the LDS-ring X-engines, the kernels the benchmark times.

* k_xengine_ring<F16> / <F32>: exactly 128 inputs (one diagonal tile) with a
  multiple of KT spectra per call (32 FP16, 16 FP32) -- the mid-size
  correlator's engine, run here through the whole F -> X chain at its shape
  (128 inputs, 8192-point FFT, FP64 phase model with 0-50 sample delays) on
  64 spectra;
* k_xengine_ring2<F16> / <F32>: a multiple of 128 inputs (several tiles per
  channel) -- the large-N engine, on 2-bit channelized voltages of 256 and 512
  inputs, ungrouped (the default) and in lockstep groups (FX_XENGINE_LOCKSTEP:
  the tiles of a channel advance together on one XCD), with channels spread
  unevenly over the groups and rounds that straddle channels.

Each is checked against the oracle (FP32 within 1e-5 of the FP32 restatement
of genericAddProduct_32fc_conj, architecture.h:427-434, and of the FP64 sums;
FP16 within 1e-3 of the FP64 oracle) and bit for bit against the
register-prefetch engine (FX_XENGINE_PREFETCH) on the same channelized data.
"""
import numpy as np
import pytest

import fxlib
import oracle as O

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


def nrel(a, b):
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _mid_inputs(T, n=8192, stations=64):
    rng = np.random.default_rng(20240229)
    tau = rng.uniform(0.0, 50.0, stations)
    models = []
    for s in range(stations):
        for _ in range(2):
            models.append((float(tau[s]), 1e-7 * s, float(rng.uniform(-np.pi, np.pi)), 1e-4))
    streams = [O.gen_station(91, i, 0.02, int(tau[i // 2]), 0.0, 0.0, 0.9816, n * T) for i in range(2 * stations)]
    return streams, models


@pytest.fixture(scope="module")
def mid64():
    """The mid-size shape on 64 spectra and its FP32 / FP64 oracle visibilities."""
    T, n = 64, 8192
    streams, models = _mid_inputs(T, n)
    ref32 = O.corr_f32(streams, n, T, models, nthreads=16)
    ref64 = O.corr_f64(streams, n, T, models, nthreads=16)
    return streams, models, ref32, ref64


@pytest.mark.parametrize("prec,name", [(fxlib.FX_PREC_F16, "k_xengine_ring<F16>"),
                                       (fxlib.FX_PREC_F32, "k_xengine_ring<F32>")])
def test_ring_engine_mid_shape(ctx, mid64, prec, name):
    streams, models, ref32, ref64 = mid64
    n, T, I = 8192, 64, len(streams)
    words = np.stack(streams)
    dw = ctx.to_device(words)
    got = {}
    for kind in (fxlib.FX_XENGINE_AUTO, fxlib.FX_XENGINE_PREFETCH):
        corr = fxlib.Correlator(ctx, I, n, T, prec, use_phase=True)
        corr.set_phase_model(models)
        corr.set_xengine(kind)
        kname = corr.xengine_name(T)
        assert kname == (name if kind == fxlib.FX_XENGINE_AUTO else name.replace("_ring", "").replace(">", ", 128>"))
        corr.process(dw, words.shape[1], 0, T)
        got[kind] = corr.visibilities()
        corr.close()
    ring, pre = got[fxlib.FX_XENGINE_AUTO], got[fxlib.FX_XENGINE_PREFETCH]
    assert np.array_equal(ring.view(np.uint32), pre.view(np.uint32))
    e64 = nrel(ring, ref64)
    if prec == fxlib.FX_PREC_F32:
        e32 = nrel(ring, ref32)
        print(f"{name}: vs FP32 oracle {e32:.2e}, vs FP64 {e64:.2e}")
        assert e32 < TOL32 and e64 < TOL32
    else:
        print(f"{name}: vs FP64 oracle {e64:.2e}")
        assert e64 < 1e-3
    # the ring engine in two calls of 32 spectra (F and X per call) is the same run
    corr = fxlib.Correlator(ctx, I, n, 32, prec, use_phase=True)
    corr.set_phase_model(models)
    corr.process(dw, words.shape[1], 0, 32)
    corr.process(dw, words.shape[1], 32, 32)
    two = corr.visibilities()
    corr.close()
    assert nrel(two, ref64) < (TOL32 if prec == fxlib.FX_PREC_F32 else 1e-3)


@pytest.mark.parametrize("I,C,T", [(256, 128, 64), (512, 128, 64), (512, 128, 256)])
@pytest.mark.parametrize("prec", [fxlib.FX_PREC_F32, fxlib.FX_PREC_F16])
def test_ring2_engine(ctx, I, C, T, prec):
    codes = ctx.alloc(C * I * T)
    ctx.synth_channelized(7 + I + C, C, T, I, 0.01, 0.9816, codes, input_major=False)
    esz = 8 if prec == fxlib.FX_PREC_F32 else 4
    dx = ctx.alloc(C * T * I * esz)
    ctx.codes_to_chan(codes, False, C, T, I, prec, dx)
    p = "F32" if prec == fxlib.FX_PREC_F32 else "F16"
    names = {fxlib.FX_XENGINE_AUTO: f"k_xengine_ring2<{p}>",
             fxlib.FX_XENGINE_LOCKSTEP: f"k_xengine_ring2<{p}> (lockstep)",
             fxlib.FX_XENGINE_PREFETCH: f"k_xengine<{p}, 128>"}
    got = {}
    for kind, want in names.items():
        corr = fxlib.Correlator(ctx, I, 2 * C, 8, prec)
        corr.set_scales(np.ones(C, np.float32), np.ones(I, np.float32))
        corr.set_xengine(kind)
        assert corr.xengine_name(T, dx) == want
        corr.xengine(dx, T)
        got[kind] = corr.visibilities()
        corr.close()
    ring, pre = got[fxlib.FX_XENGINE_AUTO], got[fxlib.FX_XENGINE_PREFETCH]
    assert np.array_equal(ring.view(np.uint32), pre.view(np.uint32))
    assert np.array_equal(got[fxlib.FX_XENGINE_LOCKSTEP].view(np.uint32), pre.view(np.uint32))
    cb = codes.download(np.uint8).reshape(C, T, I)
    lut = O.LUT.astype(np.float64)
    X = lut[cb & 3] + 1j * lut[(cb >> 2) & 3]              # [c][t][i], level-table values
    iu = np.triu_indices(I)
    ref64 = np.stack([(X[c].T @ X[c].conj())[iu] for c in range(C)], 1)   # FP64 sums
    e64 = nrel(ring, ref64)
    if prec == fxlib.FX_PREC_F32:
        sample = [0, C // 2, C - 1]
        ref32 = O.xengine_f32(X[sample].astype(np.complex64))   # the FP32 restatement
        e32 = nrel(ring[:, sample], ref32)
        print(f"k_xengine_ring2<F32> {I} inputs: vs FP32 oracle {e32:.2e}, vs FP64 {e64:.2e}")
        assert e32 < TOL32 and e64 < TOL32
    else:
        print(f"k_xengine_ring2<F16> {I} inputs: vs FP64 {e64:.2e}")
        # the level values +-1 and +-3.3359 are FP16 numbers to within 2^-11
        assert e64 < 1e-3
    codes.free()
    dx.free()

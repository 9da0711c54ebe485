"""GPU parity of the batched uneven-subband path (fx_batch_*), the 64 x 64
X-engine tile and the autotuner.

Each subband of a batch with uneven FFT lengths and input counts must match
the oracle: FP32 within 1e-5 of the FP32 restatement, FP16 within 1e-3 of the
FP64 oracle; the hipGraph replay must equal the launch-by-launch run exactly
on every replay (with eager work and synchronisation between replays), and
the per-input FP64 phase-cal tones must match the oracle's.
"""
import os

import numpy as np
import pytest

import bench
import fxlib
import oracle as O

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
FS = 16e6


def nrel(a, b):
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("prec", [fxlib.FX_PREC_F32, fxlib.FX_PREC_F16])
@pytest.mark.parametrize("I,C,T", [(20, 128, 24), (100, 128, 16), (64, 256, 8)])
def test_xengine_tile64(ctx, prec, I, C, T):
    rng = np.random.default_rng(I + C)
    X = (rng.standard_normal((C, T, I)) + 1j * rng.standard_normal((C, T, I))).astype(np.complex64)
    corr = fxlib.Correlator(ctx, I, 2 * C, 8, prec)
    corr.tune(0, 64)
    if prec == fxlib.FX_PREC_F16:
        corr.set_scales(np.ones(C, np.float32), np.ones(I, np.float32))
        Xh = np.empty((C, T, I, 2), np.float16)
        Xh[..., 0], Xh[..., 1] = X.real, X.imag
        dX = ctx.to_device(Xh)
    else:
        dX = ctx.to_device(X)
    corr.xengine(dX, T)
    got = corr.visibilities()
    corr.close()
    if prec == fxlib.FX_PREC_F32:
        assert nrel(got, O.xengine_f32(X)) < TOL32
    else:
        assert nrel(got, O.xengine_f64(X)) < 1e-3


def _make_batch_input(ctx):
    shapes = [(256, 16, 512), (1024, 40, 64), (4096, 24, 16), (16384, 18, 8), (1024, 96, 32)]
    subs, streams = [], []
    for k, (n, I, T) in enumerate(shapes):
        st = [O.gen_station(50 + k, i, 0.03, i % 5, 0.4 if i == 0 else 0.0, (10e3 + 1e6) / FS, 0.9816, n * T)
              for i in range(I)]
        w = np.stack(st)
        rng = np.random.default_rng(k)
        models = [(float(rng.uniform(0, 20)), 0.0, float(rng.uniform(-3, 3)), 0.0) for _ in range(I)] if k % 2 else None
        subs.append({"nfft": n, "ninp": I, "nspectra": T, "words": ctx.to_device(w), "stride_words": w.shape[1],
                     "models": models, "pcal": (FS / 2, 1e6, 10e3) if k < 2 else None})
        streams.append(st)
    return subs, streams


def test_batch_graph_vs_eager_vs_oracle(ctx, tmp_path):
    subs, streams = _make_batch_input(ctx)
    b32 = fxlib.Batch(ctx, subs, fxlib.FX_PREC_F32, use_graph=False)
    b32.run()
    b32.sync()
    tune = str(tmp_path / "tune.txt")
    g16 = fxlib.Batch(ctx, subs, fxlib.FX_PREC_F16, use_graph=True, tune_file=tune)
    e16 = fxlib.Batch(ctx, subs, fxlib.FX_PREC_F16, use_graph=False, tune_file=tune)
    assert os.path.exists(tune)
    lines = [ln for ln in open(tune) if ln.strip() and not ln.startswith("#")]
    assert len(lines) == len({(s["nfft"], s["ninp"]) for s in subs})
    e16.run()
    e16.sync()
    ref16 = [e16.visibilities(k) for k in range(len(subs))]
    for rep in range(3):
        g16.run()
        g16.sync()
        for k in range(len(subs)):
            assert np.array_equal(g16.visibilities(k), ref16[k]), (rep, k)
        # eager work between the replays
        e16.run()
        e16.sync()
    for k, s in enumerate(subs):
        n, I, T = s["nfft"], s["ninp"], s["nspectra"]
        r32 = O.corr_f32(streams[k], n, T, s["models"], nthreads=8)
        r64 = O.corr_f64(streams[k], n, T, s["models"], nthreads=8)
        assert nrel(b32.visibilities(k), r32) < TOL32, k
        assert nrel(ref16[k], r64) < 1e-3, k
        sub, xt, chunk = g16.tuning(k)
        assert xt in (64, 128) and sub % 8 == 0 and chunk >= T
        if s["pcal"] is not None:
            cfg = O.pcal_config(*s["pcal"])
            for i in (0, I - 1):
                x = O.unpack_1ch(streams[k][i], n * T)
                ref = O.pcal_final(cfg, O.pcal_extract(cfg, x))
                for bb in (b32, g16):
                    assert nrel(bb.pcal(k, i), ref) < 1e-12
    for bb in (b32, g16, e16):
        bb.close()


def test_autotune_cache(ctx, tmp_path):
    tune = str(tmp_path / "t.txt")
    import ctypes as C
    sub, xt = C.c_long(), C.c_int()
    fxlib.check(fxlib.load().fx_autotune(ctx.handle, 1024, 40, fxlib.FX_PREC_F16, 2048, tune.encode(),
                                         C.byref(sub), C.byref(xt)))
    # the sub-chunk candidates that fit a 2048-spectrum chunk
    assert sub.value in (64, 256, 1024) and xt.value in (64, 128)
    f = open(tune).read().split()
    assert f[:6] == ["1024", "40", str(fxlib.FX_PREC_F16), "2048", str(sub.value), str(xt.value)]


def test_batch_seeded_draw_all_subbands(ctx, tmp_path):
    """The benchmark's seeded draw of 32 uneven subbands (FFT lengths from
    {256, 1024, 4096, 16384}, 16-96 inputs, rho = 0.03, a phase-cal tone on
    input 0) on a shortened integration of 2^17 samples per input: every
    subband of the FP16 graph within 1e-3 of the FP64 oracle, of the FP32
    batch within 1e-5 of the FP32 oracle, phase-cal tones within 1e-12."""
    nsamp = 1 << 17
    lib = fxlib.load()
    subs, streams = [], []
    for k, (n, I, T) in enumerate(bench.batch_draw(20240229 + 4, 32, nsamp)):
        nw = nsamp // 16
        w = ctx.alloc(I * nw * 4)
        for i in range(I):
            fxlib.check(lib.fx_synth_station(ctx.handle, 20240229 + 100 + k, i, 0.03, i % 7, 0.2 if i == 0 else 0.0,
                                             (10e3 + 3e6) / FS, 0.9816, nsamp, w.ptr.value + i * nw * 4))
        subs.append({"nfft": n, "ninp": I, "nspectra": T, "words": w, "stride_words": nw, "models": None,
                     "pcal": (FS / 2, 1e6, 10e3)})
    ctx.sync()
    for s in subs:
        streams.append([s["words"].download_at(i * s["stride_words"] * 4, np.uint32, s["stride_words"])
                        for i in range(s["ninp"])])
    g16 = fxlib.Batch(ctx, subs, fxlib.FX_PREC_F16, use_graph=True, tune_file=str(tmp_path / "tune.txt"))
    b32 = fxlib.Batch(ctx, subs, fxlib.FX_PREC_F32, use_graph=False)
    for b in (g16, b32):
        b.run()
        b.sync()
    worst16 = worst32 = 0.0
    for k, s in enumerate(subs):
        n, I, T = s["nfft"], s["ninp"], s["nspectra"]
        r64 = O.corr_f64(streams[k], n, T, None, nthreads=16)
        r32 = O.corr_f32(streams[k], n, T, None, nthreads=16)
        e16, e32 = nrel(g16.visibilities(k), r64), nrel(b32.visibilities(k), r32)
        worst16, worst32 = max(worst16, e16), max(worst32, e32)
        assert e16 < 1e-3, (k, n, I, e16)
        assert e32 < TOL32, (k, n, I, e32)
        cfg = O.pcal_config(*s["pcal"])
        x = O.unpack_1ch(streams[k][0], n * T)
        ref = O.pcal_final(cfg, O.pcal_extract(cfg, x))
        for b in (g16, b32):
            assert nrel(b.pcal(k, 0), ref) < 1e-12
    print(f"32-subband draw: worst FP16 vs FP64 {worst16:.2e}, worst FP32 vs FP32 oracle {worst32:.2e}")
    g16.close()
    b32.close()
    for s in subs:
        s["words"].free()


@pytest.mark.parametrize("mode", [fxlib.FX_BATCH_PERSISTENT, fxlib.FX_BATCH_PHASED])
def test_batch_persistent_vs_oracle(ctx, tmp_path, mode):
    """The persistent batch kernel (one launch walking the work table of every
    subband, or the same table as three phased launches): every subband's
    FP16 visibilities within 1e-3 of the FP64 oracle, its FP64 phase-cal
    tones equal to the oracle's, replays bit-identical (with eager work in
    between), and the fused and phased runs bit-identical to each other."""
    subs, streams = _make_batch_input(ctx)
    tune = str(tmp_path / "tune.txt")
    b = fxlib.Batch(ctx, subs, fxlib.FX_PREC_F16, use_graph=mode, tune_file=tune)
    assert any(ln.startswith("P ") for ln in open(tune))
    nitems, nblocks, lds = b.persist_info()
    assert nitems > 0 and nblocks >= 256 and lds <= 160 * 1024
    b.run()
    b.sync()
    first = [b.visibilities(k) for k in range(len(subs))]
    e32 = fxlib.Batch(ctx, subs, fxlib.FX_PREC_F32, use_graph=False)
    for rep in range(2):
        e32.run()
        e32.sync()
        b.run()
        b.sync()
        for k in range(len(subs)):
            assert np.array_equal(b.visibilities(k).view(np.uint32), first[k].view(np.uint32)), (rep, k)
    for k, s in enumerate(subs):
        n, I, T = s["nfft"], s["ninp"], s["nspectra"]
        r64 = O.corr_f64(streams[k], n, T, s["models"], nthreads=8)
        e = nrel(first[k], r64)
        assert e < 1e-3, (k, e)
        na, cb, g, ts, xt = b.persist_tuning(k)
        assert na in (0, 1) and cb >= 1 and g >= 1 and ts >= 1
        assert na == 0 or n >= 2048
        assert xt == 64 or (xt == 128 and 64 < I <= 128)
        if s["pcal"] is not None:
            cfg = O.pcal_config(*s["pcal"])
            for i in (0, I - 1):
                x = O.unpack_1ch(streams[k][i], n * T)
                ref = O.pcal_final(cfg, O.pcal_extract(cfg, x))
                assert nrel(b.pcal(k, i), ref) < 1e-12
    other = fxlib.Batch(ctx, subs, fxlib.FX_PREC_F16,
                        use_graph=fxlib.FX_BATCH_PHASED if mode == fxlib.FX_BATCH_PERSISTENT
                        else fxlib.FX_BATCH_PERSISTENT, tune_file=tune)
    other.run()
    other.sync()
    for k in range(len(subs)):
        assert np.array_equal(other.visibilities(k).view(np.uint32), first[k].view(np.uint32)), k
    for bb in (b, e32, other):
        bb.close()

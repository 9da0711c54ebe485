"""The three forms of the direct F kernel (FFT lengths 4096-16384) write the
same channelized samples: the persistent kernel with FP16 blocks in pairs
(the default, 32-byte store runs), the persistent kernel with single blocks,
and one block per workgroup.  Checked at the mid-size correlator's shape (128
inputs, 8192-point FFT, FP64 phase model, FP16 and FP32 out) and at the
other lengths, and against the FP32 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(ctx, fxlib, dw, stride, I, n, nt, t0, prec, models, kind, esz):
    corr = fxlib.Correlator(ctx, I, n, nt, prec, use_phase=models is not None)
    if models is not None:
        corr.set_phase_model(models)
    corr.set_fstage(kind)
    name = corr.fstage_name(nt, prec)
    dX = ctx.alloc((n // 2) * nt * I * esz)
    corr.fstage(dw, stride, t0, nt, dX, prec)
    raw = dX.download(np.uint32, shape=((n // 2) * nt * I * esz // 4,))
    dX.free()
    corr.close()
    return name, raw


@pytest.mark.parametrize("n,I,nt,phase", [(8192, 128, 32, True), (4096, 16, 32, True), (16384, 8, 16, True),
                                           (8192, 24, 16, False), (8192, 20, 16, True)])
@pytest.mark.parametrize("prec", ["f16", "f32"])
def test_fstage_forms_identical(ctx, n, I, nt, phase, prec):
    import fxlib
    import oracle as O
    from test_gpu_kernels import _streams, _upload_streams, nrel
    P = {"f16": (fxlib.FX_PREC_F16, 4), "f32": (fxlib.FX_PREC_F32, 8)}[prec]
    t0 = 3
    streams = _streams(I, n * (t0 + nt), seed=23)
    models = [(0.37 * i + 0.1, 1e-3 * i, 0.5 * i, 0.01 * i) for i in range(I)] if phase else None
    dw, stride = _upload_streams(ctx, streams)
    out = {}
    for kind in (fxlib.FX_FSTAGE_AUTO, fxlib.FX_FSTAGE_PERSISTENT, fxlib.FX_FSTAGE_BLOCKS):
        out[kind] = _run(ctx, fxlib, dw, stride, I, n, nt, t0, P[0], models, kind, P[1])
    names = {k: v[0] for k, v in out.items()}
    print(f"n={n} I={I} nt={nt} {prec}: {names}")
    assert names[fxlib.FX_FSTAGE_BLOCKS].startswith("k_fstage_direct<")
    assert names[fxlib.FX_FSTAGE_PERSISTENT].startswith("k_fstage_direct_p<")
    ref = out[fxlib.FX_FSTAGE_BLOCKS][1]
    for kind in (fxlib.FX_FSTAGE_AUTO, fxlib.FX_FSTAGE_PERSISTENT):
        assert np.array_equal(out[kind][1], ref), (names[kind], int(np.sum(out[kind][1] != ref)))
    # and the values themselves against the FP32 oracle
    if prec == "f32":
        got = ref.view(np.complex64).reshape(n // 2, nt, I)
        want = O.fstage_f32(streams, n, t0, nt, models)
        assert nrel(got, want) < 1e-5
    else:
        h = ref.view(np.float16).reshape(n // 2, nt, I, 2).astype(np.float32)
        got = h[..., 0] + 1j * h[..., 1]
        want = O.fstage_f32(streams, n, t0, nt, models) * np.float32(1.0 / np.sqrt(n * 4.316))
        assert nrel(got, want) < 1e-3
    dw.free()

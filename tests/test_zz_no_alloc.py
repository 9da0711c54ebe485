"""fx_spectrometer_ex_noalloc (the spectrometer for callers that cannot
allocate device memory) against fx_spectrometer_ex, through the C-ABI.

Two calls in a row on one GPU share its static scratch array: first the
16-channel, 32768-point whole-stream spectrum (34.6 MB of partial spectra),
then a smaller 4-channel, 1024-point overlapped run with a Hann window and
the Nyquist bin.  Both must equal the allocating entry point's spectra bit
for bit (same kernels, same launch shape, same window table).  A shape whose
partial spectra exceed the array is refused with FX_ERR_ARG before any
launch.
"""
import ctypes as C

import numpy as np
import pytest

import fxlib

pytestmark = pytest.mark.gpu


def _nspec(ntime, n, overlap, K, continuous):
    hop = n // overlap
    if continuous:
        nfft = (ntime - n) // hop + 1 if ntime >= n else 0
        return nfft // K
    return ntime // ((K - 1) * hop + n)


def _spectra(ctx, fn, dw, ntime, nch, n, ov, K, win, cont, nyq):
    lib = fxlib.load()
    ns = _nspec(ntime, n, ov, K, cont)
    nbins = n // 2 + (1 if nyq else 0)
    out = ctx.alloc(max(1, ns) * nch * nbins * 4)
    # every element must be written: start from a NaN pattern
    fxlib.check(lib.fx_memset(ctx.handle, out.ptr, 0xFF, max(1, ns) * nch * nbins * 4))
    got = C.c_long(-1)
    fxlib.check(getattr(lib, fn)(ctx.handle, dw.ptr, ntime, nch, n, ov, K, win, cont, nyq, out.ptr, C.byref(got)))
    assert got.value == ns
    ctx.sync()
    v = out.download(np.float32, shape=(ns, nch, nbins))
    out.free()
    return v


@pytest.mark.parametrize("case", ["two_calls"])
def test_noalloc_spectrometer_matches_allocating_entry(case):
    with fxlib.Context(0) as ctx:
        # call 1: 16 interleaved channels, 2^24 samples each, 32768-point FFTs,
        # one spectrum over the whole stream (SIN2 window, Nyquist kept)
        nch, n = 16, 32768
        ntime = 1 << 24
        rng = np.random.default_rng(5)
        amp, freq = np.zeros(nch), np.zeros(nch)
        for c in rng.choice(nch, 3, replace=False):
            amp[c], freq[c] = 0.2, rng.uniform(0.05, 0.45)
        dw = ctx.alloc(ntime * nch // 4)
        ctx.synth_multichannel(17, nch, amp, freq, 0.9816, ntime, dw)
        args = (ntime, nch, n, 1, ntime // n, fxlib.FX_WIN_SIN2, 0, 1)
        ref = _spectra(ctx, "fx_spectrometer_ex", dw, *args)
        got = _spectra(ctx, "fx_spectrometer_ex_noalloc", dw, *args)
        assert np.isfinite(ref).all()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        dw.free()
        # call 2, smaller: 4 channels, 1024-point FFTs at 50% overlap, Hann,
        # 37 FFTs per spectrum, one unbroken run, a ragged sample count
        nch, n = 4, 1024
        ntime = 37 * 512 * 5 + 700
        dw = ctx.alloc((ntime * nch + 15) // 16 * 4)
        ctx.synth_multichannel(18, nch, np.zeros(nch), np.zeros(nch), 0.9816, ntime, dw)
        args = (ntime, nch, n, 2, 37, fxlib.FX_WIN_HANN, 1, 1)
        ref = _spectra(ctx, "fx_spectrometer_ex", dw, *args)
        got = _spectra(ctx, "fx_spectrometer_ex_noalloc", dw, *args)
        assert ref.shape[0] == 5 and np.isfinite(ref).all()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        # beyond the static array: refused before any launch
        lib = fxlib.load()
        big_t = 1 << 26
        out = ctx.alloc(4)
        ns = C.c_long(-1)
        rc = lib.fx_spectrometer_ex_noalloc(ctx.handle, dw.ptr, big_t, 16, 1024, 1, 1, fxlib.FX_WIN_NONE, 0, 0,
                                            out.ptr, C.byref(ns))
        assert rc == -1 and b"static scratch" in lib.fx_last_error()
        out.free()
        dw.free()

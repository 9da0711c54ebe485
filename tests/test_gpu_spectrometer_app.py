"""GPU parity of the spectrometer application: the cross spectrum
(TaskCoreIPP.cpp:456-487) against the FP64 oracle, and the dispatcher loop
(TaskDispatcher.cpp:38-230) from files to binary sinks against the oracle
run over the whole file.

Bar: spectra and cross spectra <= 1e-5 normwise (FP32 FFTs against FP64 or
the FP32 restatement); phase-cal <= 1e-12 (FP64 fold of exact 2-bit values).
"""
import os

import numpy as np
import pytest

import frames as FR
import oracle as O
import tasklib as TL

pytestmark = pytest.mark.gpu


def nrel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.mark.parametrize("n,ov,K,window,continuous", [(1024, 2, 6, O.WIN_HANN, False),
                                                      (256, 1, 9, O.WIN_NONE, True),
                                                      (8192, 4, 3, O.WIN_SIN2, False)])
def test_xpol_vs_oracle(ctx, n, ov, K, window, continuous):
    rng = np.random.default_rng(n + K)
    ntime = n * 60
    wa = rng.integers(0, 2**32, ntime * 2 // 16, dtype=np.uint64).astype(np.uint32)   # 2 channels
    wb = rng.integers(0, 2**32, ntime * 4 // 16, dtype=np.uint64).astype(np.uint32)   # 4 channels
    da, db = ctx.to_device(wa), ctx.to_device(wb)
    ref = O.spectrometer_xpol_f64(wa, 2, 1, wb, 4, 2, ntime, n, ov, K, window, continuous)
    dout = ctx.alloc(max(1, ref.shape[0]) * (n // 2 + 1) * 8)
    ns = ctx.spectrometer_xpol(da, 2, 1, db, 4, 2, ntime, n, ov, K, window, continuous, dout)
    assert ns == ref.shape[0] > 0
    got = dout.download(np.complex64, ns * (n // 2 + 1)).reshape(ns, n // 2 + 1)
    assert nrel(got, ref) < 1e-5
    assert np.all(got[:, 0].imag == 0) and np.all(got[:, -1].imag == 0)
    # the channel with itself is the auto spectrum of the spectrometer
    auto = O.spectrometer_ex_f32(wa, ntime, 2, n, ov, K, window, continuous, True)[:, 1, :]
    ns = ctx.spectrometer_xpol(da, 2, 1, da, 2, 1, ntime, n, ov, K, window, continuous, dout)
    got = dout.download(np.complex64, ns * (n // 2 + 1)).reshape(ns, n // 2 + 1)
    assert nrel(got.real, auto) < 1e-5


def _settings(n, ov, K, window, nch, fmt, base, xpol=0, pcal=0, fs=16e6, off=0.0, harm=1e6, chunk=3,
              chans=(1, 1), framebytes=0, fps=0):
    s = TL.SwspecSettings()
    c = s.core
    c.fft_points, c.fft_overlap_factor, c.core_overlapped_ffts = n, ov, K
    c.window, c.bits_per_sample, c.source_channels = window, 2, nch
    c.extract_pcal, c.samplingfreq, c.pcaloffsethz, c.pcalharmonicshz = pcal, fs, off, harm
    s.format = fmt
    s.framebytes, s.headerbytes, s.frames_per_second = framebytes, 32, fps
    s.use_channel_file[0], s.use_channel_file[1] = chans
    s.calc_xpol, s.spectra_per_chunk = xpol, chunk
    s.basefilename[0] = (base + "_a").encode()
    s.basefilename[1] = (base + "_b").encode()
    return s


def _read(path, dtype, width):
    x = np.fromfile(path, dtype=dtype)
    return x.reshape(-1, width)


@pytest.mark.parametrize("chunk", [1, 3, 7])
def test_app_two_raw_files_with_xpol(tmp_path, chunk):
    import fxlib
    n, ov, K, nch = 1024, 2, 5, 2
    L = (K - 1) * (n // ov) + n
    ntime = L * 11 + 704                      # a partial spectrum at the end
    rng = np.random.default_rng(chunk)
    w = [rng.integers(0, 2**32, ntime * nch // 16, dtype=np.uint64).astype(np.uint32) for _ in range(2)]
    paths = [str(tmp_path / f"in{i}.raw") for i in range(2)]
    for p, x in zip(paths, w):
        x.tofile(p)
    base = str(tmp_path / "out")
    s = _settings(n, ov, K, O.WIN_HANN, nch, fxlib.FX_FMT_RAW, base, xpol=1, chunk=chunk, chans=(2, 1))
    st = TL.SwspecStats()
    lib = TL.load()
    assert lib.fx_swspectrometer_run(s, paths[0].encode(), paths[1].encode(), st) == 0, \
        fxlib.load().fx_last_error()
    assert st.spectra == 11
    C1 = n // 2 + 1
    a = _read(base + "_a_swspec.bin", np.float32, C1)
    b = _read(base + "_b_swspec.bin", np.float32, C1)
    x = _read(base + "_a_xpol_swspec.bin", np.complex64, C1)
    assert a.shape == b.shape == x.shape == (11, C1)
    ra = O.spectrometer_ex_f32(w[0], ntime, nch, n, ov, K, O.WIN_HANN, False, True)[:, 1, :]
    rb = O.spectrometer_ex_f32(w[1], ntime, nch, n, ov, K, O.WIN_HANN, False, True)[:, 0, :]
    assert nrel(a, ra) < 1e-5 and nrel(b, rb) < 1e-5
    rx = O.spectrometer_xpol_f64(w[0], nch, 1, w[1], nch, 0, ntime, n, ov, K, O.WIN_HANN)
    assert nrel(x, rx) < 1e-5


def test_app_vdif_with_pcal(tmp_path):
    """One single-channel VDIF file (8032-byte frames in order) with a
    phase-cal comb: auto spectra and per-spectrum phase-cal against the
    oracle over the concatenated payloads."""
    import fxlib
    n, ov, K = 1024, 2, 5
    fs, off, harm = 16e6, 500e3, 1e6
    T, R = 16, 32
    pb = 8000
    fps = int(fs * 2 / 8 / pb)                # 2-bit samples, one channel: 4000 frames/s
    nfr = 90
    ntime = nfr * pb * 4
    words = O.gen_multichannel(21, 1, [0.3], [(off + 3 * harm) / fs], 0.9816, ntime)
    payload = words.view(np.uint8)
    frames = []
    for k in range(nfr):
        hdr = FR.vdif_encode(1000 + k // fps, k % fps, pb + 32, thread=3, epoch=40)
        frames.append(hdr + payload[k * pb:(k + 1) * pb].tobytes())
    path = str(tmp_path / "obs.vdif")
    with open(path, "wb") as fh:
        fh.write(b"".join(frames))
    base = str(tmp_path / "vd")
    s = _settings(n, ov, K, O.WIN_SIN2, 1, fxlib.FX_FMT_VDIF, base, pcal=1, fs=fs, off=off, harm=harm, chunk=4,
                  framebytes=pb + 32, fps=fps)
    st = TL.SwspecStats()
    assert TL.load().fx_swspectrometer_run(s, path.encode(), None, st) == 0, fxlib.load().fx_last_error()
    L = (K - 1) * (n // ov) + n
    nspec = ntime // L
    assert st.spectra == nspec
    spec = _read(base + "_a_swspec.bin", np.float32, n // 2 + 1)
    ref = O.spectrometer_ex_f32(words, ntime, 1, n, ov, K, O.WIN_SIN2, False, True)[:, 0, :]
    assert spec.shape == ref.shape and nrel(spec, ref) < 1e-5
    pc = _read(base + "_a_pcal.bin", np.complex64, T)
    x = O.unpack_1ch(words, ntime)
    hop = n // ov
    for sp in range(nspec):
        acc = np.zeros(T, np.complex128)
        for j in range(K):
            if (sp * K + j) % ov == 0:
                st0 = sp * L + j * hop
                O.tcpcal_fold(x[st0: st0 + n], T, R, -2 * np.pi * off / fs, acc)
        # the sink holds float32 of the FP64 result
        assert nrel(pc[sp], acc.astype(np.complex64)) < 1e-6


def test_app_vdif_blanked_frames(tmp_path):
    """Invalid-bit and fill-pattern VDIF frames decode as zeros in the spectra
    and the phase-cal (mark5access blanking), against the oracle with the
    same words zeroed."""
    import fxlib
    n, ov, K = 1024, 2, 5
    fs, off, harm = 16e6, 500e3, 1e6
    T, R = 16, 32
    pb, nfr, fps = 8000, 60, 4000
    ntime = nfr * pb * 4
    words = O.gen_multichannel(23, 1, [0.3], [(off + 3 * harm) / fs], 0.9816, ntime)
    payload = bytearray(words.view(np.uint8).tobytes())
    bad_bit, bad_fill = {7, 30}, {44}
    for k in bad_fill:                          # fill at the end of the payload
        payload[(k + 1) * pb - 8:(k + 1) * pb] = FR.FILL64.to_bytes(8, "little")
    frames = []
    for k in range(nfr):
        hdr = FR.vdif_encode(1000, k, pb + 32, thread=0, epoch=40, invalid=k in bad_bit)
        frames.append(hdr + bytes(payload[k * pb:(k + 1) * pb]))
    path = str(tmp_path / "bad.vdif")
    with open(path, "wb") as fh:
        fh.write(b"".join(frames))
    base = str(tmp_path / "bl")
    s = _settings(n, ov, K, O.WIN_HANN, 1, fxlib.FX_FMT_VDIF, base, pcal=1, fs=fs, off=off, harm=harm, chunk=9,
                  framebytes=pb + 32, fps=fps)
    st = TL.SwspecStats()
    assert TL.load().fx_swspectrometer_run(s, path.encode(), None, st) == 0, fxlib.load().fx_last_error()
    placed = np.frombuffer(bytes(payload), dtype="<u4")
    valid = np.ones(placed.size, bool)
    for k in bad_bit | bad_fill:
        valid[k * pb // 4:(k + 1) * pb // 4] = False
    L = (K - 1) * (n // ov) + n
    nspec = ntime // L
    assert st.spectra == nspec
    spec = _read(base + "_a_swspec.bin", np.float32, n // 2 + 1)
    ref = O.spectrometer_xpol_f64(placed, 1, 0, placed, 1, 0, ntime, n, ov, K, O.WIN_HANN, valid_a=valid,
                                  valid_b=valid).real
    assert spec.shape == ref.shape and nrel(spec, ref) < 1e-5
    # spectra that touch a blanked frame differ from the unblanked ones
    plain = O.spectrometer_ex_f32(placed, ntime, 1, n, ov, K, O.WIN_HANN, False, True)[:, 0, :]
    touched = [sp for sp in range(nspec) if not valid[sp * L // 16:(sp * L + L) // 16].all()]
    assert touched and all(nrel(spec[sp], plain[sp]) > 1e-3 for sp in touched)
    pc = _read(base + "_a_pcal.bin", np.complex64, T)
    x = O.unpack_1ch(placed, ntime)
    x[~valid[np.arange(ntime) // 16]] = 0.0
    hop = n // ov
    for sp in range(nspec):
        acc = np.zeros(T, np.complex128)
        for j in range(K):
            if (sp * K + j) % ov == 0:
                st0 = sp * L + j * hop
                O.tcpcal_fold(x[st0: st0 + n], T, R, -2 * np.pi * off / fs, acc)
        assert nrel(pc[sp], acc.astype(np.complex64)) < 1e-6


def test_app_mark5b(tmp_path):
    """A Mark5B file (10016-byte frames, sign/magnitude sample codes) crossing
    a second boundary: the spectra equal the oracle's over the samples the
    frames carry."""
    import fxlib
    n, ov, K = 1024, 2, 5
    fps, nfr = 400, 40                         # 16 MHz, 2-bit, one channel: 400 frames/s
    ntime = nfr * 10000 * 4
    words = O.gen_multichannel(33, 1, [0.2], [0.1], 0.9816, ntime)
    m5 = FR.mark5b_swap_pairs(words).astype("<u4").view(np.uint8)   # the swap is its own inverse
    frames = []
    for k in range(nfr):
        f = fps - 25 + k                      # 25 frames of second 86398, the rest of 86399
        frames.append(FR.mark5b_encode(417, 86398 + f // fps, f % fps) + m5[k * 10000:(k + 1) * 10000].tobytes())
    path = str(tmp_path / "obs.m5b")
    with open(path, "wb") as fh:
        fh.write(b"".join(frames))
    base = str(tmp_path / "m5")
    s = _settings(n, ov, K, O.WIN_HANN, 1, fxlib.FX_FMT_MARK5B, base, chunk=5, framebytes=10016, fps=fps)
    s.headerbytes = 16
    st = TL.SwspecStats()
    assert TL.load().fx_swspectrometer_run(s, path.encode(), None, st) == 0, fxlib.load().fx_last_error()
    L = (K - 1) * (n // ov) + n
    assert st.spectra == ntime // L
    spec = _read(base + "_a_swspec.bin", np.float32, n // 2 + 1)
    ref = O.spectrometer_ex_f32(words, ntime, 1, n, ov, K, O.WIN_HANN, False, True)[:, 0, :]
    assert spec.shape == ref.shape and nrel(spec, ref) < 1e-5


def test_app_rejects_bad_settings(tmp_path):
    import fxlib
    p = str(tmp_path / "x.raw")
    np.zeros(1024, np.uint32).tofile(p)
    lib = TL.load()
    for kw in (dict(nch=3), dict(chans=(5, 1)), dict(fmt=fxlib.FX_FMT_MARK5B), dict(fmt=7), dict(chunk=0)):
        s = _settings(1024, 2, 4, O.WIN_NONE, kw.get("nch", 2), kw.get("fmt", fxlib.FX_FMT_RAW),
                      str(tmp_path / "o"), chunk=kw.get("chunk", 2), chans=kw.get("chans", (1, 1)))
        assert lib.fx_swspectrometer_run(s, p.encode(), None, None) != 0
    s = _settings(1024, 2, 4, O.WIN_NONE, 2, fxlib.FX_FMT_RAW, str(tmp_path / "o"))
    assert lib.fx_swspectrometer_run(s, str(tmp_path / "missing.raw").encode(), None, None) != 0

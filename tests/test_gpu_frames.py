"""GPU parity of the real-file front end: fx_frames_place over VDIF and Mark5B
streams (format_vdif.c:135-206,2299-2345; mark5_format_mark5b.c:134-186;
blanker_mark5.c:54-200,282-320) against the oracle, and the double-buffered
file ingest against the file's bytes.

Bar: bit-exact placed payload words, valid ranges and counters.
"""
import os

import numpy as np
import pytest

import frames as FR

pytestmark = pytest.mark.gpu


def _rand_words(rng, n):
    return rng.integers(0, 1 << 32, n, dtype=np.uint64).astype("<u4")


def _vdif_frames(rng, threads, fps, nsec, pbytes, sec0, legacy=False):
    hb = 16 if legacy else 32
    frames = []
    for s in range(nsec):
        for f in range(fps):
            for t in threads:
                hdr = FR.vdif_encode(sec0 + s, f, pbytes + hb, thread=t, epoch=42, nchan=4, legacy=legacy)
                frames.append(bytearray(hdr + _rand_words(rng, pbytes // 4).tobytes()))
    return frames


def _place_both(ctx, raw, nframes, framebytes, hb, fmt, thread_row, nrows, sec0, frame0, fps, nslots,
                stride, exact=False):
    import fxlib
    d_raw = ctx.to_device(np.frombuffer(bytes(raw), dtype=np.uint8))
    d_map = ctx.to_device(np.asarray(thread_row, dtype=np.int32))
    d_out = ctx.alloc(4 * nrows * stride)
    d_out.zero()
    d_valid = ctx.alloc(8 * nrows * nslots)
    d_counts = ctx.alloc(8 * 4)
    ctx.frames_place(fmt, d_raw, nframes, framebytes, hb, d_map, len(thread_row), nrows, sec0, frame0, fps,
                     nslots, d_out, stride, d_valid, d_counts, clear=True, exact_blank=exact)
    got = (d_out.download(np.uint32).reshape(nrows, stride),
           d_valid.download(np.uint32).reshape(nrows, nslots, 2),
           d_counts.download(np.uint64))
    want = FR.frames_place(bytes(raw), nframes, framebytes, hb, fmt, thread_row, nrows, sec0, frame0, fps,
                           nslots, stride, exact=exact)
    assert fxlib is not None
    return got, want


def _assert_same(got, want):
    for g, w, name in zip(got, want, ("payload", "valid", "counts")):
        assert np.array_equal(g, w), name


@pytest.mark.parametrize("legacy", [False, True])
def test_vdif_place_shuffled_threads(ctx, legacy):
    import fxlib
    rng = np.random.default_rng(11 + legacy)
    threads, fps, nsec, pbytes, sec0 = [3, 9, 0, 17], 8, 3, 1024, 123456
    hb = 16 if legacy else 32
    frames = _vdif_frames(rng, threads, fps, nsec, pbytes, sec0, legacy)
    n = len(frames)
    # damage: invalid bit, fill at either end, one fill word only (still
    # valid), frames outside the window, an unwanted thread
    frames[5][3] |= 0x80
    frames[9][hb:hb + 8] = FR.FILL64.to_bytes(8, "little")
    frames[14][-8:] = FR.FILL64.to_bytes(8, "little")
    frames[20][-4:] = FR.FILL32.to_bytes(4, "little")
    frames.append(bytearray(FR.vdif_encode(sec0 + nsec, 0, pbytes + hb, thread=3, legacy=legacy)
                            + bytes(pbytes)))
    frames.append(bytearray(FR.vdif_encode(sec0 - 1, fps - 1, pbytes + hb, thread=9, legacy=legacy)
                            + bytes(pbytes)))
    frames.append(bytearray(FR.vdif_encode(sec0, 1, pbytes + hb, thread=5, legacy=legacy) + bytes(pbytes)))
    drop = set(rng.choice(n, 6, replace=False).tolist()) - {5, 9, 14, 20}
    frames = [f for i, f in enumerate(frames) if i not in drop]
    order = rng.permutation(len(frames))
    raw = b"".join(bytes(frames[i]) for i in order)
    thread_row = [-1] * 32
    for r, t in enumerate(threads):
        thread_row[t] = r
    nslots = fps * nsec
    stride = nslots * pbytes // 4 + 64
    got, want = _place_both(ctx, raw, len(frames), pbytes + hb, hb, fxlib.FX_FMT_VDIF, thread_row,
                            len(threads), sec0, 0, fps, nslots, stride)
    _assert_same(got, want)
    counts = got[2]
    assert counts[0] == n - len(drop) and counts[1] == 3 and counts[2] == 2 and counts[3] == 1
    # missing frames leave their slot's range empty
    assert int(np.sum(got[1][..., 1] > got[1][..., 0])) == n - len(drop) - 3


def _mark5b_frames(rng, nframes, day, sod0, fps, frame0=0):
    frames = []
    for k in range(nframes):
        f = frame0 + k
        hdr = FR.mark5b_encode(day, sod0 + f // fps, f % fps)
        frames.append(bytearray(hdr + _rand_words(rng, 2500).tobytes()))
    return frames


@pytest.mark.parametrize("exact", [False, True])
def test_mark5b_place_partial_fill(ctx, exact):
    import fxlib
    rng = np.random.default_rng(21 + exact)
    fps, nframes, day, sod0 = 6400, 40, 417, 86390
    frames = _mark5b_frames(rng, nframes, day, sod0, fps, frame0=6390)   # crosses a second
    fill = FR.FILL32.to_bytes(4, "little")
    for k in range(1, 30, 2):
        b = int(rng.integers(1, 2500)) * 4            # fill boundary in bytes
        if k % 4 == 1:
            frames[k][16 + b:] = fill * (2500 - b // 4)     # trailing fill
        else:
            frames[k][16:16 + b] = fill * (b // 4)          # leading fill
    frames[30][16:] = fill * 2500                      # all fill
    frames[31][0:4] = (0).to_bytes(4, "little")        # no sync word
    frames[32][16 + 8:16 + 16] = FR.FILL64.to_bytes(8, "little")   # fill inside: valid
    raw = b"".join(bytes(frames[i]) for i in rng.permutation(nframes))
    sec0 = day * 86400 + sod0
    nslots = 48
    got, want = _place_both(ctx, raw, nframes, 10016, 16, fxlib.FX_FMT_MARK5B, [0], 1, sec0, 6390 - 6, fps,
                            nslots, nslots * 2500, exact=exact)
    _assert_same(got, want)
    assert got[2][0] == nframes and got[2][1] >= 2


def test_mark5b_payload_codes_swapped(ctx):
    """A placed Mark5B payload unpacks to the values mark5access decodes."""
    import fxlib
    rng = np.random.default_rng(5)
    frames = _mark5b_frames(rng, 1, 100, 5, 6400)
    raw = bytes(frames[0])
    d_raw = ctx.to_device(np.frombuffer(raw, dtype=np.uint8))
    d_map = ctx.to_device(np.zeros(1, np.int32))
    d_out, d_valid, d_counts = ctx.alloc(10000), ctx.alloc(8), ctx.alloc(32)
    ctx.frames_place(fxlib.FX_FMT_MARK5B, d_raw, 1, 10016, 16, d_map, 1, 1, 100 * 86400 + 5, 0, 6400, 1,
                     d_out, 2500, d_valid, d_counts)
    d_f = ctx.alloc(4 * 40000)
    ctx.unpack2b(d_out, 40000, 1, d_f)
    got = d_f.download(np.float32)
    # mark5access 2-bit decode of the raw payload: per byte, 4 samples, code =
    # bit 2i (sign) + 2 * bit 2i+1 (magnitude) -> {-H, +1, -1, +H}
    lut = np.array([-3.3359, 1.0, -1.0, 3.3359], np.float32)
    b = np.frombuffer(raw[16:], np.uint8)
    codes = np.stack([((b >> (2 * i)) & 1) + 2 * ((b >> (2 * i + 1)) & 1) for i in range(4)], axis=1)
    assert np.array_equal(got, lut[codes.ravel()])


def test_frames_place_accumulates_and_rejects(ctx):
    import fxlib
    rng = np.random.default_rng(6)
    frames = _vdif_frames(rng, [0, 1], 4, 1, 256, 50)
    raw = b"".join(bytes(f) for f in frames)
    d_raw = ctx.to_device(np.frombuffer(raw, dtype=np.uint8))
    d_map = ctx.to_device(np.array([0, 1], np.int32))
    stride = 4 * 64
    d_out, d_valid, d_counts = ctx.alloc(4 * 2 * stride), ctx.alloc(8 * 2 * 4), ctx.alloc(32)
    fb = 256 + 32
    # two calls over halves of the stream, the second without clearing
    ctx.frames_place(fxlib.FX_FMT_VDIF, d_raw, 4, fb, 32, d_map, 2, 2, 50, 0, 4, 4, d_out, stride, d_valid,
                     d_counts, clear=True)
    ctx.frames_place(fxlib.FX_FMT_VDIF, d_raw, 4, fb, 32, d_map, 2, 2, 50, 0, 4, 4, d_out, stride, d_valid,
                     d_counts, clear=False, raw_offset=4 * fb)
    want = FR.frames_place(raw, 8, fb, 32, FR.FMT_VDIF, [0, 1], 2, 50, 0, 4, 4, stride)
    _assert_same((d_out.download(np.uint32).reshape(2, stride), d_valid.download(np.uint32).reshape(2, 4, 2),
                  d_counts.download(np.uint64)), want)
    ctx.frames_place(fxlib.FX_FMT_VDIF, d_raw, 0, fb, 32, d_map, 2, 2, 50, 0, 4, 4, d_out, stride, d_valid,
                     d_counts, clear=True)                       # no frames: only the clear
    assert not d_counts.download(np.uint64).any()
    bad = [dict(framebytes=fb + 4), dict(hb=24), dict(fmt=7), dict(stride=stride - 1), dict(fps=0)]
    for b in bad:
        with pytest.raises(fxlib.FxError):
            ctx.frames_place(b.get("fmt", fxlib.FX_FMT_VDIF), d_raw, 8, b.get("framebytes", fb),
                             b.get("hb", 32), d_map, 2, 2, 50, 0, b.get("fps", 4), 4, d_out,
                             b.get("stride", stride), d_valid, d_counts)
    with pytest.raises(fxlib.FxError):                          # Mark5B needs 10016-byte frames
        ctx.frames_place(fxlib.FX_FMT_MARK5B, d_raw, 1, fb, 16, d_map, 1, 1, 0, 0, 1, 1, d_out, stride,
                         d_valid, d_counts)


@pytest.mark.parametrize("chunk,nbuf,offset", [(1 << 20, 2, 0), (777_777, 3, 100), (64 << 10, 2, 5)])
def test_ingest_stream_bytes(ctx, tmp_path, chunk, nbuf, offset):
    import fxlib
    rng = np.random.default_rng(chunk)
    data = rng.integers(0, 256, 5_300_000, dtype=np.uint8)
    path = os.path.join(tmp_path, "stream.bin")
    data.tofile(path)
    g = fxlib.Ingest(ctx, path, chunk, nbuf=nbuf, offset=offset)
    pos = offset
    got_chunks = 0
    for p, n in g:
        host = np.empty(n, np.uint8)
        fxlib.check(fxlib.load().fx_memcpy_d2h(ctx.handle, host.ctypes.data_as(fxlib.C.c_void_p),
                                               fxlib.C.c_void_p(p), n))
        assert n == min(chunk, data.size - pos)
        assert np.array_equal(host, data[pos:pos + n])
        pos += n
        got_chunks += 1
    assert pos == data.size and got_chunks == -(-(data.size - offset) // chunk)
    assert g.next() == (None, 0)
    g.close()
    with pytest.raises(fxlib.FxError):
        fxlib.Ingest(ctx, os.path.join(tmp_path, "missing.bin"), chunk)


def test_ingest_then_place_vdif(ctx, tmp_path):
    """File -> double-buffered chunks of whole frames -> placement per chunk,
    equal to the oracle over the whole file."""
    import fxlib
    rng = np.random.default_rng(8)
    threads, fps, nsec, pbytes, sec0 = [0, 1, 2], 16, 2, 2048, 7000
    frames = _vdif_frames(rng, threads, fps, nsec, pbytes, sec0)
    order = rng.permutation(len(frames))
    raw = b"".join(bytes(frames[i]) for i in order)
    path = os.path.join(tmp_path, "obs.vdif")
    with open(path, "wb") as fh:
        fh.write(raw)
    fb = pbytes + 32
    nslots = fps * nsec
    stride = nslots * pbytes // 4
    d_map = ctx.to_device(np.array(threads, np.int32))
    d_out, d_valid, d_counts = ctx.alloc(4 * 3 * stride), ctx.alloc(8 * 3 * nslots), ctx.alloc(32)
    d_out.zero()
    first = True
    g = fxlib.Ingest(ctx, path, 7 * fb, nbuf=2)
    for p, n in g:
        assert n % fb == 0
        ctx.frames_place(fxlib.FX_FMT_VDIF, p, n // fb, fb, 32, d_map, 3, 3, sec0, 0, fps, nslots, d_out,
                         stride, d_valid, d_counts, clear=first)
        first = False
    g.close()
    want = FR.frames_place(raw, len(frames), fb, 32, FR.FMT_VDIF, threads, 3, sec0, 0, fps, nslots, stride)
    _assert_same((d_out.download(np.uint32).reshape(3, stride), d_valid.download(np.uint32).reshape(3, nslots, 2),
                  d_counts.download(np.uint64)), want)


def test_vdif_file_to_visibilities(ctx, tmp_path):
    """A 4-thread VDIF file (one station stream per thread, frames shuffled)
    streamed through the ingest, placed per thread and correlated: the
    visibilities equal the FP32 oracle's on the original streams."""
    import fxlib
    import oracle as O
    n, nspec, ninp = 1024, 64, 4
    nsamp = n * nspec                             # 65536 samples = 16 KiB per input
    streams = [O.gen_station(77, i, 0.05, i, 0.0, 0.0, 0.9816, nsamp) for i in range(ninp)]
    pb = 8192                                     # payload bytes: 2 frames per thread
    frames = []
    for t, w in enumerate(streams):
        b = w.astype("<u4").tobytes()
        for f in range(len(b) // pb):
            frames.append(FR.vdif_encode(500, f, pb + 32, thread=10 + t) + b[f * pb:(f + 1) * pb])
    rng = np.random.default_rng(9)
    raw = b"".join(frames[i] for i in rng.permutation(len(frames)))
    path = str(tmp_path / "stations.vdif")
    with open(path, "wb") as fh:
        fh.write(raw)
    fb, pw, nslots = pb + 32, pb // 4, 2
    stride = nslots * pw
    thread_row = [-1] * 16
    for t in range(ninp):
        thread_row[10 + t] = t
    d_map = ctx.to_device(np.array(thread_row, np.int32))
    d_rows, d_valid, d_counts = ctx.alloc(4 * ninp * stride), ctx.alloc(8 * ninp * nslots), ctx.alloc(32)
    g = fxlib.Ingest(ctx, path, 3 * fb, nbuf=2)
    first = True
    for p, nb in g:
        ctx.frames_place(fxlib.FX_FMT_VDIF, p, nb // fb, fb, 32, d_map, 16, ninp, 500, 0, 1000, nslots, d_rows,
                         stride, d_valid, d_counts, clear=first)
        first = False
    g.close()
    counts = d_counts.download(np.uint64)
    assert list(counts) == [len(frames), 0, 0, 0]
    corr = fxlib.Correlator(ctx, ninp, n, nspec, fxlib.FX_PREC_F32)
    corr.process(d_rows, stride, 0, nspec)
    got = corr.visibilities()
    corr.close()
    ref = O.corr_f32(streams, n, nspec)
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    assert err < 1e-5, err

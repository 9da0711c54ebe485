"""The front-end oracle (oracle/frames.py) pinned, and the host-side header
decoders of the library checked against it (no GPU: fx_vdif_decode and
fx_mark5b_decode are plain host functions).

Pins: MJDs of the VDIF reference epochs against dates the mark5access table
lists (format_vdif.c:143-153) and the Jan/Jul day counts; hand-assembled VDIF
and Mark5B headers with fields worked out from the header layouts; the exact
blanker against a linear scan for every fill boundary of a Mark5B payload;
the mark5access bisection against a probe sequence worked by hand; the Mark5B
sample-code swap against the two lookup tables (mark5_format_mark5b.c:59-96).
"""
import numpy as np
import pytest

import frames as FR

H = 3.3359


def test_vdif_epoch_mjd_known_dates():
    # dates from the mark5access epoch table: 2000-01-01, 2000-07-01,
    # 2004-07-01, 2008-07-01, 2020-01-01, 2024-01-01, 2031-07-01
    known = {0: 51544, 1: 51726, 9: 53187, 17: 54648, 40: 58849, 48: 60310, 63: 63048}
    for e, mjd in known.items():
        assert FR.vdif_epoch_mjd(e) == mjd


def test_vdif_epoch_day_counts():
    # 1 Jan -> 1 Jul is 181 days (182 in a leap year); 1 Jul -> 1 Jan is 184
    for e in range(63):
        d = FR.vdif_epoch_mjd(e + 1) - FR.vdif_epoch_mjd(e)
        year = 2000 + e // 2
        leap = year % 4 == 0 and (year % 100 != 0 or year % 400 == 0)
        assert d == ((182 if leap else 181) if e % 2 == 0 else 184)


def _bytes(words):
    return np.array(words, dtype="<u4").tobytes()


# seconds 74565, epoch 48 (2024-01-01), frame 1234, version 3, 4 channels,
# 8032-byte frames (1004 x 8), 2 bits, thread 5, station "AB"
VDIF_KNOWN = _bytes([0x00012345, 0x300004D2, 0x620003EC, 0x04054142, 0, 0, 0, 0])
VDIF_KNOWN_FIELDS = dict(seconds=74565, legacy=0, invalid=0, frame=1234, epoch=48, framebytes=8032,
                         headerbytes=32, nchan=4, version=3, station=0x4142, thread=5, nbits=2,
                         is_complex=0, mjd=60310, second_of_day=74565)


def test_vdif_known_header():
    assert FR.vdif_decode(VDIF_KNOWN) == VDIF_KNOWN_FIELDS
    legacy_invalid = _bytes([0xC0000000 | (3 * 86400 + 7), (1 << 24) | 9, 0x00000003, 0x83FF0000])
    h = FR.vdif_decode(legacy_invalid)
    assert (h["legacy"], h["invalid"], h["headerbytes"], h["framebytes"]) == (1, 1, 16, 24)
    assert (h["mjd"], h["second_of_day"], h["frame"]) == (51726 + 3, 7, 9)
    assert (h["thread"], h["nbits"], h["is_complex"], h["nchan"]) == (1023, 1, 1, 1)


def test_vdif_encode_roundtrip():
    rng = np.random.default_rng(1)
    for _ in range(300):
        f = dict(seconds=int(rng.integers(0, 1 << 30)), frame=int(rng.integers(0, 1 << 24)),
                 framebytes=8 * int(rng.integers(3, 1 << 20)), epoch=int(rng.integers(0, 64)),
                 thread=int(rng.integers(0, 1024)), station=int(rng.integers(0, 1 << 16)),
                 nchan=1 << int(rng.integers(0, 10)), nbits=int(rng.integers(1, 33)),
                 legacy=bool(rng.integers(0, 2)), invalid=bool(rng.integers(0, 2)),
                 version=int(rng.integers(0, 8)), is_complex=bool(rng.integers(0, 2)))
        h = FR.vdif_decode(FR.vdif_encode(**f))
        for k in ("seconds", "frame", "framebytes", "epoch", "thread", "station", "nchan", "nbits",
                  "version"):
            assert h[k] == f[k], k
        assert (h["legacy"], h["invalid"], h["is_complex"]) == (f["legacy"], f["invalid"], f["is_complex"])


# sync, frame 0x1234 = 4660, day 123, second 45678, fraction 0.1234 s
# (1234 x 0.1 ms) unrounded up to the 156.25 us grid: 790 x 156250 ns
MARK5B_KNOWN = _bytes([0xABADDEED, 0x00001234, 0x12345678, 0x12340000])


def test_mark5b_known_header():
    h = FR.mark5b_decode(MARK5B_KNOWN)
    assert h == dict(sync_ok=1, frame=4660, tvg=0, day=123, second_of_day=45678, frac_ns=123437500.0)
    h = FR.mark5b_decode(_bytes([0xABADDEEC, 0x0000FFFF, 0x99986399, 0x00000000]))
    assert (h["sync_ok"], h["frame"], h["tvg"], h["day"], h["second_of_day"], h["frac_ns"]) == \
        (0, 0x7FFF, 1, 999, 86399, 0.0)


def test_mark5b_encode_roundtrip():
    rng = np.random.default_rng(2)
    for _ in range(300):
        day, sod, fr = int(rng.integers(0, 1000)), int(rng.integers(0, 86400)), int(rng.integers(0, 32768))
        tick = int(rng.integers(0, 6400))          # frames of 156.25 us in a second
        h = FR.mark5b_decode(FR.mark5b_encode(day, sod, fr, frac_ns=tick * 156250))
        assert (h["day"], h["second_of_day"], h["frame"], h["sync_ok"]) == (day, sod, fr, 1)
        # 0.1 ms truncation in the header, unrounded back to the frame grid
        assert h["frac_ns"] == tick * 156250


def test_library_decoders_match_oracle():
    import fxlib
    fxlib.load()
    assert fxlib.vdif_decode(VDIF_KNOWN) == VDIF_KNOWN_FIELDS
    rng = np.random.default_rng(3)
    for _ in range(2000):
        b = rng.integers(0, 1 << 32, 4, dtype=np.uint64).astype("<u4").tobytes()
        assert fxlib.vdif_decode(b) == FR.vdif_decode(b)
        assert fxlib.mark5b_decode(b) == FR.mark5b_decode(b)
    assert fxlib.mark5b_decode(MARK5B_KNOWN) == FR.mark5b_decode(MARK5B_KNOWN)


def _fill_payload(n64, lo_bytes=None, hi_bytes=None):
    """Payload of n64 data words with fill in [0, lo_bytes) or [hi_bytes, end)."""
    w = np.full(2 * n64, 0xDEADBEEF, dtype="<u4")
    if lo_bytes is not None:
        w[:lo_bytes // 4] = FR.FILL32
    if hi_bytes is not None:
        w[hi_bytes // 4:] = FR.FILL32
    return w


def test_exact_blanker_matches_linear_scan():
    n64 = 1250                               # a Mark5B payload
    for e in range(8, 8 * n64, 4):           # fill from byte e to the end
        w = _fill_payload(n64, hi_bytes=e)
        d64 = w.view("<u8")
        if d64[-1] != FR.FILL64:
            continue
        assert FR.first_invalid_byte(d64, w, exact=True) == e
        assert FR.blank_range(w.tobytes(), FR.FMT_MARK5B, True, exact=True) == (0, e // 4)
    for b in range(8, 8 * n64 - 8, 4):       # fill from the start to byte b
        w = _fill_payload(n64, lo_bytes=b)
        d64 = w.view("<u8")
        assert FR.first_valid_byte(d64, w, exact=True) == b
        assert FR.blank_range(w.tobytes(), FR.FMT_MARK5B, True, exact=True) == (b // 4, 2 * n64)


def test_reference_bisection_known_probes():
    # fill from byte 5600 (64-bit word 700) of a 1250-word payload.  Probes:
    # 625 data -> 937 fill -> 781 fill -> 703 fill -> 664 -> 683 -> 692 ->
    # 696 -> 698 data; the landing index 699 is never probed, r = 703 -> 5624
    # bytes, and the 32-bit word before is fill -> 5620
    w = _fill_payload(1250, hi_bytes=5600)
    assert FR.first_invalid_byte(w.view("<u8"), w) == 5620
    # leading fill to byte 5600: probes 625 fill -> 937 data -> 781 data ->
    # 703 data -> 664 fill -> 683 -> 692 -> 696 -> 698 fill, r = 703 -> 5624;
    # the reference straggler test reads word 702's low half, data -> 5624
    w = _fill_payload(1250, lo_bytes=5600)
    assert FR.first_valid_byte(w.view("<u8"), w) == 5624
    assert FR.first_valid_byte(w.view("<u8"), w, exact=True) == 5600


def test_reference_straggler_defect():
    # where the mark5access bisection lands on an 8-byte-aligned boundary b
    # exactly, its straggler test reads a fill word and reports b + 4
    hits = 0
    for b in range(48, 8 * 1250 - 48, 8):
        w = _fill_payload(1250, lo_bytes=b)
        d64 = w.view("<u8")
        if 8 * FR._bisect_ref(d64, False) == b:
            hits += 1
            assert FR.first_valid_byte(d64, w) == b + 4
            assert FR.first_valid_byte(d64, w, exact=True) == b
    assert hits > 100


def test_reference_bisection_error_bounds():
    n64 = 1250
    errs = []
    for e in range(40, 8 * n64 - 40, 4):
        w = _fill_payload(n64, hi_bytes=e)
        errs.append(FR.first_invalid_byte(w.view("<u8"), w) - e)
    errs = np.array(errs)
    assert errs.min() == 0 and errs.max() <= 36 and np.all(errs % 4 == 0)
    # exact where the last probe lands on the boundary: most cases
    assert np.mean(errs == 0) > 0.4


def test_vdif_blank_all_or_nothing():
    w = np.arange(2000, dtype="<u4") + 7
    assert FR.blank_range(w.tobytes(), FR.FMT_VDIF, True) == (0, 2000)
    assert FR.blank_range(w.tobytes(), FR.FMT_VDIF, False) == (0, 0)
    w2 = w.copy(); w2[-2:] = FR.FILL32
    assert FR.blank_range(w2.tobytes(), FR.FMT_VDIF, True) == (0, 0)
    w3 = w.copy(); w3[-1] = FR.FILL32      # one fill word: not the 64-bit pattern
    assert FR.blank_range(w3.tobytes(), FR.FMT_VDIF, True) == (0, 2000)


def test_mark5b_code_swap_matches_luts():
    ours = np.array([-H, -1.0, 1.0, H])
    mark5b = np.array([-H, 1.0, -1.0, H])   # index: sign bit + 2 * magnitude bit
    codes = np.arange(1 << 16, dtype=np.uint32)
    sw = FR.mark5b_swap_pairs(codes)
    for j in range(8):
        c = (codes >> (2 * j)) & 3
        s = (sw >> np.uint32(2 * j)) & 3
        assert np.array_equal(ours[s], mark5b[c])


def _vdif_stream(rng, threads, fps, nsec, pbytes, sec_base=1000):
    frames, keys = [], []
    for s in range(nsec):
        for f in range(fps):
            for t in threads:
                hdr = FR.vdif_encode(sec_base + s, f, pbytes + 32, thread=t, epoch=40)
                pay = rng.integers(0, 1 << 32, pbytes // 4, dtype=np.uint64).astype("<u4").tobytes()
                frames.append(hdr + pay)
                keys.append((t, s * fps + f))
    return frames, keys


def test_frames_place_oracle_placement_and_permutation():
    rng = np.random.default_rng(4)
    threads, fps, nsec, pbytes = [3, 7, 1], 5, 2, 64
    frames, keys = _vdif_stream(rng, threads, fps, nsec, pbytes)
    thread_row = [-1] * 8
    for r, t in enumerate(threads):
        thread_row[t] = r
    nslots, pw = fps * nsec, pbytes // 4
    args = (len(frames), pbytes + 32, 32, FR.FMT_VDIF, thread_row, len(threads), 1000, 0, fps, nslots,
            nslots * pw)
    out, valid, counts = FR.frames_place(b"".join(frames), *args)
    for fr, (t, slot) in zip(frames, keys):
        r = thread_row[t]
        assert np.array_equal(out[r, slot * pw:(slot + 1) * pw], np.frombuffer(fr[32:], "<u4"))
        assert tuple(valid[r, slot]) == (0, pw)
    assert list(counts) == [len(frames), 0, 0, 0]
    perm = rng.permutation(len(frames))
    out2, valid2, counts2 = FR.frames_place(b"".join(frames[i] for i in perm), *args)
    assert np.array_equal(out, out2) and np.array_equal(valid, valid2) and np.array_equal(counts, counts2)

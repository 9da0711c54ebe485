"""CPU restatement of the real-file front end: VDIF and Mark5B frames.

TEST INFRASTRUCTURE ONLY: the checker for fx_vdif_decode, fx_mark5b_decode
and fx_frames_place.  Imported by tests/ only; never by product code.

Pure Python and numpy over small streams (tens of frames).  Each function
cites the mark5access source it restates (spectrometer/mark5access/
mark5access/...).  The encoders build synthetic streams for the tests; they
follow the VDIF and Mark5B header layouts field by field and are pinned by a
hand-assembled known-answer header and decode round trips.
"""
from __future__ import annotations

import datetime

import numpy as np

FMT_VDIF, FMT_MARK5B = 0, 1
FILL32 = 0x11223344
FILL64 = 0x1122334411223344
MARK5B_SYNC = 0xABADDEED
MARK5B_FRAMEBYTES = 10016
MARK5B_HEADERBYTES = 16


def mjd_of(year, month, day):
    """Modified Julian Date of a Gregorian date (MJD 0 = 1858-11-17)."""
    return datetime.date(year, month, day).toordinal() - datetime.date(1858, 11, 17).toordinal()


def vdif_epoch_mjd(epoch):
    """MJD of VDIF reference epoch e: 1 January (e even) or 1 July (e odd) of
    2000 + e/2 (the mjdepochs table, format_vdif.c:143-153)."""
    return mjd_of(2000 + epoch // 2, 7 if epoch & 1 else 1, 1)


def _le32(b, off):
    return int(b[off]) | int(b[off + 1]) << 8 | int(b[off + 2]) << 16 | int(b[off + 3]) << 24


def vdif_decode(b):
    """Header fields of a VDIF frame (format_vdif.c:155-206: seconds and epoch
    to MJD and second of day, frame number; the remaining fields per the VDIF
    header layout)."""
    w0, w1, w2, w3 = (_le32(b, 4 * k) for k in range(4))
    seconds = w0 & 0x3FFFFFFF
    legacy = (w0 >> 30) & 1
    epoch = (w1 >> 24) & 0x3F
    return {
        "seconds": seconds,
        "legacy": legacy,
        "invalid": (w0 >> 31) & 1,
        "frame": w1 & 0xFFFFFF,
        "epoch": epoch,
        "framebytes": (w2 & 0xFFFFFF) * 8,
        "headerbytes": 16 if legacy else 32,
        "nchan": 1 << ((w2 >> 24) & 0x1F),
        "version": (w2 >> 29) & 7,
        "station": w3 & 0xFFFF,
        "thread": (w3 >> 16) & 0x3FF,
        "nbits": ((w3 >> 26) & 0x1F) + 1,
        "is_complex": (w3 >> 31) & 1,
        "mjd": vdif_epoch_mjd(epoch) + seconds // 86400,
        "second_of_day": seconds % 86400,
    }


def vdif_encode(seconds, frame, framebytes, *, epoch=0, thread=0, station=0, nchan=1, nbits=2,
                legacy=False, invalid=False, version=0, is_complex=False):
    """A VDIF header (16 bytes if legacy, else 32 with zero extended words)."""
    log2nchan = int(nchan).bit_length() - 1
    assert 1 << log2nchan == nchan and framebytes % 8 == 0
    w = [0] * (4 if legacy else 8)
    w[0] = (seconds & 0x3FFFFFFF) | (int(legacy) << 30) | (int(invalid) << 31)
    w[1] = (frame & 0xFFFFFF) | ((epoch & 0x3F) << 24)
    w[2] = ((framebytes // 8) & 0xFFFFFF) | (log2nchan << 24) | ((version & 7) << 29)
    w[3] = (station & 0xFFFF) | ((thread & 0x3FF) << 16) | (((nbits - 1) & 0x1F) << 26) | (int(is_complex) << 31)
    return np.array(w, dtype="<u4").tobytes()


def _bcd(value, ndigits):
    """Digits of value as BCD nibbles, most significant first."""
    return int(f"{value:0{ndigits}d}", 16)


def mark5b_encode(day, second_of_day, frame, *, frac_ns=None, tvg=False, sync=True):
    """A Mark5B header (mark5_format_mark5b.c:134-186 read it back): sync word,
    frame number, BCD JJJSSSSS, BCD fraction in 0.1 ms units."""
    if frac_ns is None:
        frac_ns = 0
    frac_01ms = int(frac_ns) // 100000
    w0 = MARK5B_SYNC if sync else 0
    w1 = (frame & 0x7FFF) | (int(tvg) << 15)
    w2 = _bcd(day % 1000, 3) << 20 | _bcd(second_of_day, 5)
    w3 = _bcd(frac_01ms, 4) << 16
    return np.array([w0, w1, w2, w3], dtype="<u4").tobytes()


def mark5b_decode(b):
    """Header fields of a Mark5B frame (mark5_format_mark5b.c:134-186): nibbles
    of header word 2 (bytes 8-11, little-endian) give day (MJD mod 1000) and
    second of day; the top four nibbles of word 3 the fraction in 0.1 ms,
    "unrounded" up to the 156.25 us frame grid."""
    w0, w1, w2, w3 = (_le32(b, 4 * k) for k in range(4))
    nib = [(w2 >> (28 - 4 * k)) & 0xF for k in range(8)] + [(w3 >> (28 - 4 * k)) & 0xF for k in range(8)]
    ns = nib[8] * 100000000 + nib[9] * 10000000 + nib[10] * 1000000 + nib[11] * 100000
    return {
        "sync_ok": int(w0 == MARK5B_SYNC),
        "frame": w1 & 0x7FFF,
        "tvg": (w1 >> 15) & 1,
        "day": nib[0] * 100 + nib[1] * 10 + nib[2],
        "second_of_day": nib[3] * 10000 + nib[4] * 1000 + nib[5] * 100 + nib[6] * 10 + nib[7],
        "frac_ns": float(156250 * ((ns + 156249) // 156250)),
    }


def _bisect_ref(d64, want_fill):
    """The mark5access bisection (blanker_mark5.c:65-77, 97-109): start at n/2,
    steps n/4, n/8, ..., 1; r is the last probed index whose word is (or, for
    want_fill False, is not) fill, n-1 if none.  The index the last step lands
    on is never probed, so r can lie a few words past the boundary."""
    n = len(d64)
    i, r, s = n // 2, n - 1, n // 4
    while s > 0:
        assert 0 <= i < n
        hit = int(d64[i]) == FILL64
        if hit == want_fill:
            r = i
            i -= s
        else:
            i += s
        s >>= 1
    return r


def _bisect_exact(d64, want_fill):
    """First index of the run at the end ([0] is not in it, [n-1] is)."""
    lo, hi = 0, len(d64) - 1
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if (int(d64[mid]) == FILL64) == want_fill:
            hi = mid
        else:
            lo = mid
    return hi


def first_invalid_byte(d64, d32, exact=False):
    """findfirstinvalid (blanker_mark5.c:54-85), one zone starting at word 0:
    start of a trailing fill run in bytes, then 4 bytes back when the 32-bit
    word before it is fill."""
    r = 8 * (_bisect_exact(d64, True) if exact else _bisect_ref(d64, True))
    assert r >= 8
    if int(d32[r // 4 - 1]) == FILL32:
        r -= 4
    return r


def first_valid_byte(d64, d32, exact=False):
    """findfirstvalid (blanker_mark5.c:87-118), one zone starting at word 0:
    end of a leading fill run in bytes, then 4 bytes on when the first 32-bit
    half of the first data word is still fill.  The reference tests
    d32[r/4 - 2] (the low half of the word before, fill whenever the search
    is exact), restated as such by default; exact=True bisects exactly and
    tests d32[r/4]."""
    if exact:
        r = 8 * _bisect_exact(d64, False)
        if int(d32[r // 4]) == FILL32:
            r += 4
        return r
    r = 8 * _bisect_ref(d64, False)
    assert r >= 8
    if int(d32[r // 4 - 2]) == FILL32:
        r += 4
    return r


def blank_range(payload, fmt, header_ok, exact=False):
    """Valid 32-bit word range [start, end) of one payload (bytes).

    VDIF (blanker_mark5.c:282-320, format_vdif.c:2328-2335): all or nothing;
    nothing with the invalid bit or fill at either end.  Mark5B
    (blanker_mark5.c:122-200, one zone as a 10000-byte payload is below the
    32 KB zone size): both ends valid -> all, both fill -> nothing, else the
    bisected boundary; nothing without the sync word."""
    d64 = np.frombuffer(payload, dtype="<u8")
    d32 = np.frombuffer(payload, dtype="<u4")
    pwords = len(d32)
    if not header_ok:
        return 0, 0
    start_ok, end_ok = int(d64[0]) != FILL64, int(d64[-1]) != FILL64
    if fmt == FMT_VDIF:
        return (0, pwords) if start_ok and end_ok else (0, 0)
    if start_ok and end_ok:
        return 0, pwords
    if not start_ok and not end_ok:
        return 0, 0
    if start_ok:
        return 0, first_invalid_byte(d64, d32, exact) // 4
    return first_valid_byte(d64, d32, exact) // 4, pwords


def mark5b_swap_pairs(words):
    """Mark5B 2-bit code (sign bit low, magnitude bit high: -H, +1, -1, +H,
    mark5_format_mark5b.c:59-96) to this library's code (-H, -1, +1, +H)."""
    w = np.asarray(words, dtype=np.uint32)
    return ((w & np.uint32(0x55555555)) << np.uint32(1)) | ((w >> np.uint32(1)) & np.uint32(0x55555555))


def frames_place(raw, nframes, framebytes, payload_offset, fmt, thread_row, nrows, sec0, frame0, fps,
                 nslots, out_stride, out=None, valid=None, counts=None, exact=False):
    """Frame-by-frame restatement of fx_frames_place: decode, map thread to
    row, time to slot, blank, copy.  Frames are assumed to hit distinct slots.
    Returns (out words [nrows, out_stride], valid [nrows, nslots, 2], counts)."""
    pwords = (framebytes - payload_offset) // 4
    if out is None:
        out = np.zeros((nrows, out_stride), dtype=np.uint32)
    if valid is None:
        valid = np.zeros((nrows, nslots, 2), dtype=np.uint32)
    if counts is None:
        counts = np.zeros(4, dtype=np.uint64)
    raw = bytes(raw)
    for fr in range(nframes):
        f = raw[fr * framebytes:(fr + 1) * framebytes]
        if fmt == FMT_VDIF:
            h = vdif_decode(f)
            sec, fnum, thread, ok = h["seconds"], h["frame"], h["thread"], not h["invalid"]
        else:
            h = mark5b_decode(f)
            sec = h["day"] * 86400 + h["second_of_day"]
            fnum, thread, ok = h["frame"], 0, bool(h["sync_ok"])
        row = thread_row[thread] if thread < len(thread_row) else -1
        if row >= nrows:
            row = -1
        slot = (sec - sec0) * fps + fnum - frame0
        if row < 0:
            counts[3] += 1
            continue
        if slot < 0 or slot >= nslots:
            counts[2] += 1
            continue
        payload = f[payload_offset:]
        vs, ve = blank_range(payload, fmt, ok, exact)
        valid[row, slot] = (vs, ve)
        counts[0] += 1
        if ve <= vs:
            counts[1] += 1
        words = np.frombuffer(payload, dtype="<u4")
        if fmt == FMT_MARK5B:
            words = mark5b_swap_pairs(words)
        out[row, slot * pwords:(slot + 1) * pwords] = words
    return out, valid, counts

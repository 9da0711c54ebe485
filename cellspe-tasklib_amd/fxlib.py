"""ctypes binding of libcellspe_fx.so (include/cellspe_fx.h and the task-library
headers).

This is the Python-side view of the C-ABI used by the tests and the benchmark.
It holds no compute of its own: every call goes to the gfx950 kernels through
the shared library, and importing it fails loudly when the library has not
been built.
"""
from __future__ import annotations

import ctypes as C
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CELLSPE_FX_LIB names another build of the library (the sanitizer build)
LIB_PATH = os.environ.get("CELLSPE_FX_LIB") or os.path.join(HERE, "libcellspe_fx.so")

FX_OK = 0
FX_PREC_F32, FX_PREC_F16, FX_PREC_FP8, FX_PREC_FP8_2BIT, FX_PREC_FP4_2BIT = 0, 1, 2, 3, 4
FX_XENGINE_AUTO, FX_XENGINE_PREFETCH, FX_XENGINE_LOCKSTEP = 0, 1, 2
FX_FSTAGE_AUTO, FX_FSTAGE_PERSISTENT, FX_FSTAGE_BLOCKS = 0, 1, 2
FX_BATCH_EAGER, FX_BATCH_GRAPH, FX_BATCH_PERSISTENT, FX_BATCH_PHASED = 0, 1, 2, 3
(FX_WIN_NONE, FX_WIN_SIN2, FX_WIN_COSINE, FX_WIN_COSINE2, FX_WIN_HAMMING, FX_WIN_BLACKMAN,
 FX_WIN_HANN) = range(7)


class FxError(RuntimeError):
    pass


class PhaseModel(C.Structure):
    _fields_ = [("delay0", C.c_double), ("delay_rate", C.c_double),
                ("fringe0", C.c_double), ("fringe_rate", C.c_double)]


class CorrConfig(C.Structure):
    _fields_ = [("ninp", C.c_int), ("nfft", C.c_int), ("chunk_spectra", C.c_long),
                ("prec", C.c_int), ("use_phase", C.c_int)]


FX_FMT_VDIF, FX_FMT_MARK5B, FX_FMT_RAW = 0, 1, 2
FX_PLACE_CLEAR, FX_PLACE_EXACT_BLANK = 1, 2


class VdifHeader(C.Structure):
    _fields_ = [("seconds", C.c_uint32), ("legacy", C.c_int), ("invalid", C.c_int),
                ("frame", C.c_uint32), ("epoch", C.c_int), ("framebytes", C.c_int),
                ("headerbytes", C.c_int), ("nchan", C.c_uint32), ("version", C.c_int),
                ("station", C.c_int), ("thread", C.c_int), ("nbits", C.c_int),
                ("is_complex", C.c_int), ("mjd", C.c_int), ("second_of_day", C.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Mark5bHeader(C.Structure):
    _fields_ = [("sync_ok", C.c_int), ("frame", C.c_int), ("tvg", C.c_int), ("day", C.c_int),
                ("second_of_day", C.c_int), ("frac_ns", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def load():
    """Load the shared library (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FxError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    vp, sz, i, l, d, u64 = C.c_void_p, C.c_size_t, C.c_int, C.c_long, C.c_double, C.c_uint64
    sigs = {
        "fx_last_error": (C.c_char_p, []),
        "fx_device_count": (i, []),
        "fx_device_name": (i, [i, C.c_char_p, i]),
        "fx_ctx_create": (i, [i, C.POINTER(vp)]),
        "fx_ctx_destroy": (i, [vp]),
        "fx_ctx_sync": (i, [vp]),
        "fx_ctx_stream": (vp, [vp]),
        "fx_malloc": (i, [vp, sz, C.POINTER(vp)]),
        "fx_free": (i, [vp, vp]),
        "fx_memcpy_h2d": (i, [vp, vp, vp, sz]),
        "fx_memcpy_d2h": (i, [vp, vp, vp, sz]),
        "fx_memcpy_d2d": (i, [vp, vp, vp, sz]),
        "fx_memset": (i, [vp, vp, i, sz]),
        "fx_event_create": (i, [vp, C.POINTER(vp)]),
        "fx_event_record": (i, [vp, vp]),
        "fx_event_elapsed_ms": (i, [vp, vp, C.POINTER(C.c_float)]),
        "fx_event_destroy": (i, [vp]),
        "fx_synth_station": (i, [vp, u64, i, d, l, d, d, d, sz, vp]),
        "fx_synth_multichannel": (i, [vp, u64, i, vp, vp, d, sz, vp]),
        "fx_synth_channelized": (i, [vp, u64, i, l, i, d, d, i, vp]),
        "fx_codes_to_chan": (i, [vp, vp, i, i, l, i, i, vp]),
        "fx_unpack2b": (i, [vp, vp, sz, i, vp]),
        "fx_fft_c2c": (i, [vp, vp, vp, i, l]),
        "fx_vdif_decode": (i, [vp, C.POINTER(VdifHeader)]),
        "fx_mark5b_decode": (i, [vp, C.POINTER(Mark5bHeader)]),
        "fx_frames_place": (i, [vp, i, vp, sz, i, i, vp, i, i, l, l, i, l, vp, sz, vp, vp, i]),
        "fx_ingest_open": (i, [vp, C.c_char_p, sz, sz, i, C.POINTER(vp)]),
        "fx_ingest_next": (i, [vp, C.POINTER(vp), C.POINTER(sz)]),
        "fx_ingest_close": (i, [vp]),
        "fx_spectrometer_xpol": (i, [vp, vp, i, i, vp, i, i, sz, i, i, i, i, i, vp, C.POINTER(l)]),
        "fx_lags": (i, [vp, vp, l, i, C.c_float, vp]),
        "fx_amp_phase": (i, [vp, vp, l, C.c_float, vp, vp]),
        "fx_pcal_delay": (i, [vp, vp, i, i, d, d, d, d, vp]),
        "fx_fft_r2c": (i, [vp, vp, vp, i, l, i]),
        "fx_complex_mac": (i, [vp, vp, vp, vp, l, i]),
        "fx_complex_mul": (i, [vp, vp, vp, vp, l]),
        "fx_complex_add": (i, [vp, vp, vp, vp, l]),
        "fx_copy_kernel": (i, [vp, vp, vp, sz]),
        "fx_sincos": (i, [vp, vp, vp, l, i]),
        "fx_sincos_generator": (i, [vp, d, d, vp, l]),
        "fx_spectrometer": (i, [vp, vp, sz, i, i, i, i, vp]),
        "fx_spectrometer_ex": (i, [vp, vp, sz, i, i, i, i, i, i, i, vp, C.POINTER(l)]),
        "fx_spectrometer_ex_noalloc": (i, [vp, vp, sz, i, i, i, i, i, i, i, vp, C.POINTER(l)]),
        "fx_pcal_create": (i, [vp, d, d, i, sz, C.POINTER(vp)]),
        "fx_pcal_destroy": (i, [vp]),
        "fx_pcal_clear": (i, [vp]),
        "fx_pcal_adjust_sample_offset": (i, [vp, sz]),
        "fx_pcal_extract_2bit": (i, [vp, vp, sz]),
        "fx_pcal_final": (i, [vp, vp, C.POINTER(u64)]),
        "fx_pcal_info": (i, [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i)]),
        "fx_corr_create": (i, [vp, C.POINTER(CorrConfig), C.POINTER(vp)]),
        "fx_corr_destroy": (i, [vp]),
        "fx_corr_set_phase_model": (i, [vp, vp]),
        "fx_corr_set_multipliers": (i, [vp, vp]),
        "fx_corr_set_scales": (i, [vp, vp, vp]),
        "fx_corr_process": (i, [vp, vp, sz, l, l]),
        "fx_corr_xengine": (i, [vp, vp, l]),
        "fx_corr_dump": (i, [vp, vp]),
        "fx_corr_reset": (i, [vp]),
        "fx_corr_fstage": (i, [vp, vp, sz, l, l, vp, i]),
        "fx_corr_fstage_rows": (i, [vp, vp, sz, l, l, i]),
        "fx_corr_cornerturn": (i, [vp, l, l, l, i, vp]),
        "fx_corr_fstage_depth": (l, [vp]),
        "fx_corr_fstage_launches": (i, [vp, l, i]),
        "fx_corr_tune": (i, [vp, l, i]),
        "fx_corr_set_xengine": (i, [vp, i]),
        "fx_corr_set_fstage": (i, [vp, i]),
        "fx_corr_xengine_name": (i, [vp, vp, l, C.c_char_p, i]),
        "fx_corr_fstage_name": (i, [vp, l, i, C.c_char_p, i]),
        "fx_pcal_final_async": (i, [vp]),
        "fx_pcal_tones": (i, [vp, vp, C.POINTER(u64)]),
        "fx_pcal_create_inputs": (i, [vp, d, d, i, sz, i, C.POINTER(vp)]),
        "fx_pcal_extract_2bit_inputs": (i, [vp, vp, sz, sz]),
        "fx_pcal_tones_input": (i, [vp, i, vp, C.POINTER(u64)]),
        "fx_batch_create": (i, [vp, vp, i, i, i, C.c_char_p, C.POINTER(vp)]),
        "fx_batch_run": (i, [vp]),
        "fx_batch_sync": (i, [vp]),
        "fx_batch_vis": (vp, [vp, i]),
        "fx_batch_pcal": (i, [vp, i, i, vp, C.POINTER(i)]),
        "fx_batch_tuning": (i, [vp, i, C.POINTER(l), C.POINTER(i), C.POINTER(l)]),
        "fx_batch_destroy": (i, [vp]),
        "fx_batch_persist_tuning": (i, [vp, i, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(i)]),
        "fx_batch_persist_info": (i, [vp, C.POINTER(i), C.POINTER(i), C.POINTER(sz)]),
        "fx_autotune": (i, [vp, i, i, i, l, C.c_char_p, C.POINTER(l), C.POINTER(i)]),
        "fx_corr_info": (i, [vp, C.POINTER(l), C.POINTER(sz), C.POINTER(sz)]),
        "fx_mfma_probe": (i, [vp, i, vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != FX_OK:
        msg = load().fx_last_error().decode(errors="replace")
        raise FxError(f"fx call failed ({rc}): {msg}")
    return rc


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _addr(buf):
    """Device address of a DeviceBuffer or a raw pointer value."""
    if isinstance(buf, DeviceBuffer):
        return buf.ptr.value
    if isinstance(buf, C.c_void_p):
        return buf.value
    return int(buf)


def vdif_decode(header: bytes):
    h = VdifHeader()
    check(load().fx_vdif_decode(C.c_char_p(bytes(header[:16])), C.byref(h)))
    return h.as_dict()


def mark5b_decode(header: bytes):
    h = Mark5bHeader()
    check(load().fx_mark5b_decode(C.c_char_p(bytes(header[:16])), C.byref(h)))
    return h.as_dict()


class DeviceBuffer:
    """A device allocation owned by a Context."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(load().fx_malloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p
        ctx._adopt(self)

    def free(self):
        if self.ptr is not None and self.ptr.value and self.ctx.handle is not None:
            load().fx_free(self.ctx.handle, self.ptr)
        self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(load().fx_memcpy_h2d(self.ctx.handle, self.ptr, _ptr(arr), arr.nbytes))

    def download(self, dtype, count=None, shape=None):
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        check(load().fx_memcpy_d2h(self.ctx.handle, _ptr(out), self.ptr, out.nbytes))
        return out.reshape(shape) if shape is not None else out

    def download_at(self, offset_bytes, dtype, count):
        dtype = np.dtype(dtype)
        assert offset_bytes + count * dtype.itemsize <= self.nbytes
        out = np.empty(count, dtype=dtype)
        check(load().fx_memcpy_d2h(self.ctx.handle, _ptr(out),
                                   C.c_void_p(self.ptr.value + offset_bytes), out.nbytes))
        return out

    def zero(self):
        check(load().fx_memset(self.ctx.handle, self.ptr, 0, self.nbytes))


class Context:
    def __init__(self, device=0):
        lib = load()
        h = C.c_void_p()
        check(lib.fx_ctx_create(device, C.byref(h)))
        self.handle = h
        self.lib = lib
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        """Objects holding device resources of this context: closed before it."""
        self._children.add(obj)

    def close(self):
        if self.handle is not None:
            # a buffer, correlator or batch still alive (a caller that never
            # closed it, a traceback holding it) is released while the
            # context's device and stream still exist
            for obj in list(self._children):
                obj.close()
            self.lib.fx_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        check(self.lib.fx_ctx_sync(self.handle))

    def alloc(self, nbytes):
        return DeviceBuffer(self, max(int(nbytes), 1))

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        b = self.alloc(arr.nbytes)
        b.upload(arr)
        return b

    # events on the context stream
    def event(self):
        e = C.c_void_p()
        check(self.lib.fx_event_create(self.handle, C.byref(e)))
        return e

    def record(self, ev):
        check(self.lib.fx_event_record(self.handle, ev))

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        check(self.lib.fx_event_elapsed_ms(a, b, C.byref(ms)))
        return ms.value

    # ---- primitives ---------------------------------------------------
    def unpack2b(self, d_words, ntime, nchan, d_out):
        check(self.lib.fx_unpack2b(self.handle, d_words.ptr, ntime, nchan, d_out.ptr))

    def fft_c2c(self, d_in, d_out, n, batch):
        check(self.lib.fx_fft_c2c(self.handle, d_in.ptr, d_out.ptr, n, batch))

    def frames_place(self, fmt, d_raw, nframes, framebytes, payload_offset, d_thread_row, map_size,
                     nrows, sec0, frame0, frames_per_second, nslots, d_out, out_stride_words,
                     d_valid, d_counts, clear=True, exact_blank=False, raw_offset=0):
        """Validate, strip and place VDIF/Mark5B frames into per-thread rows."""
        flags = (FX_PLACE_CLEAR if clear else 0) | (FX_PLACE_EXACT_BLANK if exact_blank else 0)
        check(self.lib.fx_frames_place(self.handle, fmt, C.c_void_p(_addr(d_raw) + raw_offset), nframes,
                                       framebytes, payload_offset, d_thread_row.ptr, map_size, nrows,
                                       sec0, frame0, frames_per_second, nslots, d_out.ptr,
                                       out_stride_words, d_valid.ptr, d_counts.ptr, flags))

    def lags(self, d_vis, nrows, nchan, scale, d_lags):
        """Lag spectra (backward FFT over channels times scale) of nrows rows."""
        check(self.lib.fx_lags(self.handle, d_vis.ptr, nrows, nchan, scale, d_lags.ptr))

    def amp_phase(self, d_z, n, scale, d_amp=None, d_phase=None):
        check(self.lib.fx_amp_phase(self.handle, d_z.ptr, n, scale, d_amp.ptr if d_amp else None,
                                    d_phase.ptr if d_phase else None))

    def pcal_delay(self, d_tones, ninp, ntones, tone0_hz, tone_step_hz, band_center_hz, band_valid_hz,
                   d_delay):
        """m5pcal delay fit (ns) of each input's tones into d_delay (ninp doubles)."""
        check(self.lib.fx_pcal_delay(self.handle, d_tones.ptr, ninp, ntones, tone0_hz, tone_step_hz,
                                     band_center_hz, band_valid_hz, d_delay.ptr))

    def fft_r2c(self, d_in, d_out, n, batch, keep_nyquist=False):
        check(self.lib.fx_fft_r2c(self.handle, d_in.ptr, d_out.ptr, n, batch, int(keep_nyquist)))

    def spectrometer(self, d_words, ntime, nchan, n, integrate_ffts, window, d_out):
        check(self.lib.fx_spectrometer(self.handle, d_words.ptr, ntime, nchan, n,
                                       integrate_ffts, window, d_out.ptr))

    def spectrometer_ex(self, d_words, ntime, nchan, n, overlap, integrate_ffts, window,
                        continuous, keep_nyquist, d_out):
        ns = C.c_long()
        check(self.lib.fx_spectrometer_ex(self.handle, d_words.ptr, ntime, nchan, n, overlap,
                                          integrate_ffts, window, int(continuous),
                                          int(keep_nyquist), d_out.ptr if d_out else None,
                                          C.byref(ns)))
        return ns.value

    def spectrometer_xpol(self, d_a, nchan_a, chan_a, d_b, nchan_b, chan_b, ntime, n, overlap, integrate_ffts,
                          window, continuous, d_out):
        """Cross spectra of two 2-bit channels; returns the spectrum count."""
        ns = C.c_long()
        check(self.lib.fx_spectrometer_xpol(self.handle, d_a.ptr, nchan_a, chan_a, d_b.ptr, nchan_b, chan_b, ntime,
                                            n, overlap, integrate_ffts, window, int(continuous), d_out.ptr,
                                            C.byref(ns)))
        return ns.value

    def complex_mac(self, a, b, acc, n, conj):
        check(self.lib.fx_complex_mac(self.handle, a.ptr, b.ptr, acc.ptr, n, int(conj)))

    def synth_station(self, seed, station, rho, delay, tone_amp, tone_freq, thresh, nsamp, d_words):
        check(self.lib.fx_synth_station(self.handle, seed, station, rho, delay, tone_amp,
                                        tone_freq, thresh, nsamp, d_words.ptr))

    def synth_multichannel(self, seed, nchan, amp, freq, thresh, ntime, d_words):
        a = np.ascontiguousarray(amp, dtype=np.float64)
        f = np.ascontiguousarray(freq, dtype=np.float64)
        check(self.lib.fx_synth_multichannel(self.handle, seed, nchan, _ptr(a), _ptr(f), thresh,
                                             ntime, d_words.ptr))

    def synth_channelized(self, seed, nchan, ntime, ninp, rho, thresh, d_codes, input_major=False):
        check(self.lib.fx_synth_channelized(self.handle, seed, nchan, ntime, ninp, rho, thresh,
                                            int(input_major), d_codes.ptr))

    def codes_to_chan(self, d_codes, input_major, nchan, ntime, ninp, out_prec, d_out):
        check(self.lib.fx_codes_to_chan(self.handle, d_codes.ptr, int(input_major), nchan, ntime, ninp,
                                        out_prec, d_out.ptr))

    def mfma_probe(self, which):
        out = np.zeros((32, 32), dtype=np.float32)
        check(self.lib.fx_mfma_probe(self.handle, which, _ptr(out)))
        return out


class Ingest:
    """Double-buffered file -> HBM stream (fx_ingest_*): iterate for
    (device pointer, bytes) chunks; each stays valid until the next one."""

    def __init__(self, ctx, path, chunk_bytes, nbuf=2, offset=0):
        self.ctx = ctx
        h = C.c_void_p()
        check(load().fx_ingest_open(ctx.handle, os.fsencode(path), offset, chunk_bytes, nbuf, C.byref(h)))
        self.handle = h
        ctx._adopt(self)

    def next(self):
        p, n = C.c_void_p(), C.c_size_t()
        check(load().fx_ingest_next(self.handle, C.byref(p), C.byref(n)))
        return (p.value, n.value) if n.value else (None, 0)

    def __iter__(self):
        while True:
            p, n = self.next()
            if not n:
                return
            yield p, n

    def close(self):
        if self.handle is not None:
            load().fx_ingest_close(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PCal:
    """Device phase-cal extractor with the reference PCal interface."""

    def __init__(self, ctx, bandwidth_hz, spacing_hz, offset_hz, sampleoffset=0, ninp=1):
        self.ctx = ctx
        self.ninp = ninp
        h = C.c_void_p()
        check(load().fx_pcal_create_inputs(ctx.handle, bandwidth_hz, spacing_hz, int(offset_hz),
                                           sampleoffset, ninp, C.byref(h)))
        self.handle = h
        ctx._adopt(self)
        k, nb, nt = C.c_int(), C.c_int(), C.c_int()
        check(load().fx_pcal_info(h, C.byref(k), C.byref(nb), C.byref(nt)))
        self.kind, self.nbins, self.ntones = k.value, nb.value, nt.value

    def clear(self):
        check(load().fx_pcal_clear(self.handle))

    def adjust_sample_offset(self, off):
        check(load().fx_pcal_adjust_sample_offset(self.handle, off))

    def extract_2bit(self, d_words, nsamp):
        check(load().fx_pcal_extract_2bit(self.handle, d_words.ptr, nsamp))

    def extract_2bit_inputs(self, d_words, stride_words, nsamp):
        """All ninp inputs in one launch; input i starts stride_words after input i-1."""
        check(load().fx_pcal_extract_2bit_inputs(self.handle, d_words.ptr, stride_words, nsamp))

    def final(self):
        """ntones complex tones (ninp x ntones for a multi-input extractor)."""
        out = np.zeros(max(self.ntones, 1) * self.ninp, dtype=np.complex128)
        n = C.c_uint64()
        check(load().fx_pcal_final(self.handle, _ptr(out), C.byref(n)))
        if self.ninp == 1:
            return out[: self.ntones], n.value
        return out.reshape(self.ninp, -1)[:, : self.ntones], n.value

    def close(self):
        if self.handle is not None:
            load().fx_pcal_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Correlator:
    def __init__(self, ctx, ninp, nfft, chunk_spectra, prec=FX_PREC_F32, use_phase=False):
        self.ctx = ctx
        self.cfg = CorrConfig(ninp, nfft, chunk_spectra, prec, int(use_phase))
        h = C.c_void_p()
        check(load().fx_corr_create(ctx.handle, C.byref(self.cfg), C.byref(h)))
        self.handle = h
        ctx._adopt(self)
        self.ninp, self.nfft, self.nchan = ninp, nfft, nfft // 2
        self.nprod = ninp * (ninp + 1) // 2

    def set_phase_model(self, models):
        arr = (PhaseModel * self.ninp)(*[PhaseModel(*m) for m in models])
        check(load().fx_corr_set_phase_model(self.handle, C.cast(arr, C.c_void_p)))

    def set_scales(self, chan=None, inp=None):
        c = None if chan is None else np.ascontiguousarray(chan, dtype=np.float32)
        i = None if inp is None else np.ascontiguousarray(inp, dtype=np.float32)
        check(load().fx_corr_set_scales(self.handle, None if c is None else _ptr(c),
                                        None if i is None else _ptr(i)))

    def process(self, d_words, stride_words, t_first, nspectra, offset_bytes=0):
        ptr = C.c_void_p(d_words.ptr.value + offset_bytes)
        check(load().fx_corr_process(self.handle, ptr, stride_words, t_first, nspectra))

    def fstage(self, d_words, stride_words, t_first, nspectra, d_chan, out_prec):
        check(load().fx_corr_fstage(self.handle, d_words.ptr, stride_words, t_first, nspectra,
                                    d_chan.ptr, out_prec))

    def fstage_rows(self, d_words, stride_words, t_first, nspectra, out_prec):
        check(load().fx_corr_fstage_rows(self.handle, d_words.ptr, stride_words, t_first, nspectra,
                                         out_prec))

    def cornerturn(self, nspectra, s0, chunk_spectra, out_prec, d_chan):
        check(load().fx_corr_cornerturn(self.handle, nspectra, s0, chunk_spectra, out_prec, d_chan.ptr))

    def fstage_depth(self):
        return load().fx_corr_fstage_depth(self.handle)

    def fstage_launches(self, nspectra, out_prec):
        """Kernel launches of fstage() for nspectra spectra (1: the direct
        engine-layout transform kernel; 2: its persistent form after the clear
        of its block counters)."""
        return load().fx_corr_fstage_launches(self.handle, nspectra, out_prec)

    def fstage_name(self, nspectra, out_prec):
        """Name of the F transform kernel fstage() runs on nspectra spectra."""
        buf = C.create_string_buffer(64)
        check(load().fx_corr_fstage_name(self.handle, nspectra, out_prec, buf, 64))
        return buf.value.decode()

    def xengine(self, d_chan, nspectra):
        check(load().fx_corr_xengine(self.handle, d_chan.ptr, nspectra))

    def dump(self, d_vis):
        check(load().fx_corr_dump(self.handle, d_vis.ptr))

    def reset(self):
        check(load().fx_corr_reset(self.handle))

    def tune(self, sub_spectra=0, xtile=0):
        check(load().fx_corr_tune(self.handle, sub_spectra, xtile))

    def set_xengine(self, kind):
        """FX_XENGINE_AUTO (LDS-ring engines where they apply),
        FX_XENGINE_LOCKSTEP (the same, the large-input one in lockstep groups)
        or FX_XENGINE_PREFETCH (the register-prefetch engine for every
        shape)."""
        check(load().fx_corr_set_xengine(self.handle, kind))

    def set_fstage(self, kind):
        """FX_FSTAGE_AUTO (persistent, FP16 blocks paired where the inputs
        allow), FX_FSTAGE_PERSISTENT (persistent, single blocks) or
        FX_FSTAGE_BLOCKS (one block per workgroup) for the direct F kernel."""
        check(load().fx_corr_set_fstage(self.handle, kind))

    def xengine_name(self, nspectra, d_chan=None):
        """Name of the X-engine kernel that runs on nspectra spectra."""
        buf = C.create_string_buffer(64)
        ptr = None if d_chan is None else d_chan.ptr
        check(load().fx_corr_xengine_name(self.handle, ptr, nspectra, buf, 64))
        return buf.value.decode()

    def visibilities(self):
        """Dump to a temporary device buffer and return [P][C] complex64."""
        b = self.ctx.alloc(self.nprod * self.nchan * 8)
        self.dump(b)
        v = b.download(np.complex64, shape=(self.nprod, self.nchan))
        b.free()
        return v

    def close(self):
        if self.handle is not None:
            load().fx_corr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SubbandDesc(C.Structure):
    _fields_ = [("nfft", C.c_int), ("ninp", C.c_int), ("nspectra", C.c_long), ("d_words", C.c_void_p),
                ("stride_words", C.c_size_t), ("pm", C.c_void_p), ("pcal", C.c_int),
                ("bandwidth_hz", C.c_double), ("pcal_spacing_hz", C.c_double), ("pcal_offset_hz", C.c_int)]


class Batch:
    """Batched subbands (fx_batch_*): one correlator per subband, optionally
    captured into one hipGraph."""

    def __init__(self, ctx, subbands, prec, use_graph=True, tune_file=None):
        """subbands: list of dicts with nfft, ninp, nspectra, words (DeviceBuffer),
        stride_words, and optional models (list of 4-tuples) and pcal
        (bandwidth_hz, spacing_hz, offset_hz)."""
        self.ctx = ctx
        n = len(subbands)
        descs = (SubbandDesc * n)()
        self._keep = []
        for k, s in enumerate(subbands):
            d = descs[k]
            d.nfft, d.ninp, d.nspectra = s["nfft"], s["ninp"], s["nspectra"]
            d.d_words = s["words"].ptr.value + s.get("offset_bytes", 0)
            d.stride_words = s["stride_words"]
            if s.get("models") is not None:
                arr = (PhaseModel * s["ninp"])(*[PhaseModel(*m) for m in s["models"]])
                self._keep.append(arr)
                d.pm = C.cast(arr, C.c_void_p)
            if s.get("pcal") is not None:
                d.pcal = 1
                d.bandwidth_hz, d.pcal_spacing_hz, d.pcal_offset_hz = s["pcal"][0], s["pcal"][1], int(s["pcal"][2])
        h = C.c_void_p()
        tf = None if tune_file is None else tune_file.encode()
        check(load().fx_batch_create(ctx.handle, C.cast(descs, C.c_void_p), n, prec, int(use_graph), tf,
                                     C.byref(h)))
        self.handle = h
        ctx._adopt(self)
        self.subbands = subbands

    def run(self):
        check(load().fx_batch_run(self.handle))

    def sync(self):
        check(load().fx_batch_sync(self.handle))

    def visibilities(self, s):
        sb = self.subbands[s]
        P = sb["ninp"] * (sb["ninp"] + 1) // 2
        C_ = sb["nfft"] // 2
        out = np.empty((P, C_), np.complex64)
        ptr = load().fx_batch_vis(self.handle, s)
        check(load().fx_memcpy_d2h(self.ctx.handle, _ptr(out), C.c_void_p(ptr), out.nbytes))
        return out

    def pcal(self, s, i):
        nt = C.c_int()
        check(load().fx_batch_pcal(self.handle, s, i, None, C.byref(nt)))
        out = np.zeros(max(nt.value, 1), np.complex128)
        check(load().fx_batch_pcal(self.handle, s, i, _ptr(out), C.byref(nt)))
        return out[: nt.value]

    def tuning(self, s):
        sub, x, ch = C.c_long(), C.c_int(), C.c_long()
        check(load().fx_batch_tuning(self.handle, s, C.byref(sub), C.byref(x), C.byref(ch)))
        return sub.value, x.value, ch.value

    def persist_tuning(self, s):
        """(natural rows + corner turns, channels per X item, F groups per
        item, X time slices, X tile width) of subband s in the persistent
        modes."""
        na, cb, g, ts, xt = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(load().fx_batch_persist_tuning(self.handle, s, C.byref(na), C.byref(cb), C.byref(g), C.byref(ts),
                                             C.byref(xt)))
        return na.value, cb.value, g.value, ts.value, xt.value

    def persist_info(self):
        """(work items, workgroups, dynamic LDS bytes) of the persistent launch."""
        n, b, l = C.c_int(), C.c_int(), C.c_size_t()
        check(load().fx_batch_persist_info(self.handle, C.byref(n), C.byref(b), C.byref(l)))
        return n.value, b.value, l.value

    def close(self):
        if self.handle is not None:
            load().fx_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

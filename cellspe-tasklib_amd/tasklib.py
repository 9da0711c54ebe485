"""ctypes view of the reference-interface seams of libcellspe_fx.so.

Three C surfaces replace the reference's entry points for the F-X path:

* the task runtime and DSP wrappers (include/celldsp_tasklib.h,
  include/celldsp_dsp.h) -- cellspe-tasklib.h:122-133 and
  cellspe-tasklib-dsp.h:27-95;
* the minicorrelator control block (include/minicorrelator_fx.h) --
  minicorrelator.h:107-119;
* the spectrometer control block and TaskCore (include/spectrometer_fx.h) --
  cell_spectrometer.h:104-115 and TaskCore.h:36-89.

The classes here only marshal numpy arrays into those calls; all work runs in
the library.  Names follow the reference (CellDSPTaskClass, TaskCore methods
prepare/run/join/finalize/combineResults).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from fxlib import FxError, load as _load_fx

# cellspe-tasklib.h:37-41
CELLDSP_TASK_INCOMPLETE = 0xFFFF
CELLDSP_TASK_OK = 0x0000
CELLDSP_TASK_GENERICERROR = 0x0001
CELLDSP_TASK_UNKOWNCOMMAND = 0x0002
CELLDSP_MAX_SPE_NR = 8
CELLDSP_MAX_NUM_IOVECTORS = 8
CDSP_BEHAV_DEFAULT = 0
CDSP_BEHAV_BLOCKING = 1
CDSP_BEHAV_GRAPH = 2

# cellspe-tasklib-dsp.h:27-46
(CELLDSP_TASK_NOP, CELLDSP_TASK_MAC, CELLDSP_TASK_MULC, CELLDSP_TASK_ADDC, CELLDSP_TASK_IOTEST,
 CELLDSP_TASK_IOBENCH, CELLDSP_TASK_FFT_INIT, CELLDSP_TASK_FFT_FREE, CELLDSP_TASK_FFT_EXEC,
 CELLDSP_TASK_FFT_EXEC_DYN, CELLDSP_TASK_SINCOS, CELLDSP_TASK_SINCOS_NORM,
 CELLDSP_TASK_SINCOS_GENERTR, CELLDSP_TASK_UNPACK, CELLDSP_TASK_RUNALL, CELLDSP_TASK_DMASIMULATION,
 CELLDSP_TASK_DIFXPROCESS, CELLDSP_TASK_SIGTERM) = range(18)

N_FFT = 1024
NUM_SPE_THREADS = 6
PPU_CORRELBUF_BYTES = N_FFT * 2 * 4


class Addr64(C.Union):
    _fields_ = [("ull", C.c_ulonglong), ("ui", C.c_uint * 2), ("p", C.c_void_p)]


class _FFT(C.Structure):
    _fields_ = [("nfft", C.c_uint), ("log2_nfft", C.c_uint)]


class _Phase(C.Structure):
    _fields_ = [("argcount", C.c_uint), ("phase_start", C.c_float), ("phase_delta", C.c_float)]


class _Difx(C.Structure):
    _fields_ = [("nfft", C.c_uint), ("log2_nfft", C.c_uint), ("useUpperOrLower", C.c_uint),
                ("twicenumchannels", C.c_uint), ("numchannels", C.c_uint)]


class _Unpack(C.Structure):
    _fields_ = [("nchan", C.c_uint)]


class Subcontext(C.Union):
    _fields_ = [("FFT", _FFT), ("Phase", _Phase), ("Difx", _Difx), ("Unpack", _Unpack)]


class CellDSPTask(C.Structure):
    """CellDSPTask_t (cellspe-tasklib.h:63-86); 128-byte alignment is applied
    by CellDSPTaskClass when it places the struct."""
    _fields_ = [
        ("command", C.c_uint),
        ("completed_flag", C.c_ushort),
        ("tick_count", C.c_uint),
        ("num_input_vectors", C.c_uint),
        ("inputvector_ptrs", Addr64 * CELLDSP_MAX_NUM_IOVECTORS),
        ("inputvector_lengths", C.c_uint * CELLDSP_MAX_NUM_IOVECTORS),
        ("num_output_vectors", C.c_uint),
        ("outputvector_ptrs", Addr64 * CELLDSP_MAX_NUM_IOVECTORS),
        ("outputvector_lengths", C.c_uint * CELLDSP_MAX_NUM_IOVECTORS),
        ("subcontext", Subcontext),
        ("next_task", Addr64),
        ("_align128", C.c_ubyte * 8),       # C pads the aligned(128) struct to 256 bytes
    ]


class MiniControlBlock(C.Structure):
    """control_block (minicorrelator.h:107-119)."""
    _fields_ = [
        ("rawdata_src", Addr64),
        ("autocorrelation_out", Addr64),
        ("baselines_out", Addr64 * 15),
        ("fmultipliers_src", Addr64),
        ("syncline", Addr64),
        ("spe_num", C.c_int),
        ("num_spes", C.c_int),
        ("ffts_total", C.c_uint),
        ("ffts_to_integrate", C.c_uint),
        ("spe_ls_listOnPPU", Addr64),
        ("spe_sig_listOnPPU", Addr64),
        ("_align128", C.c_ubyte * 72),      # aligned(128): array stride 256 bytes
    ]


class MiniStats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("spectra", C.c_ulonglong), ("dumps", C.c_uint)]


class SpecControlBlock(C.Structure):
    """control_block_t (cell_spectrometer.h:104-115)."""
    _fields_ = [
        ("rank", C.c_uint),
        ("terminate_flag", C.c_uint),
        ("fft_points", C.c_uint),
        ("fft_points_log2", C.c_uint),
        ("overlap_factor", C.c_uint),
        ("integrate_ffts", C.c_uint),
        ("samples_per_rawbyte", C.c_uint),
        ("rawif_src", Addr64),
        ("rawif_bytecount", C.c_uint),
        ("spectrum_dst", Addr64),
        ("_align128", C.c_ubyte * 72),      # aligned(128): 128 bytes
    ]


class SpecMbox(C.Structure):
    _fields_ = [("t_spu", C.c_uint), ("num_spectra", C.c_uint), ("total_ffts", C.c_uint)]


class SpectrometerSettings(C.Structure):
    """fx_spectrometer_settings: the swspect_settings_t fields the core reads."""
    _fields_ = [
        ("fft_points", C.c_int),
        ("fft_overlap_factor", C.c_int),
        ("core_overlapped_ffts", C.c_int),
        ("window", C.c_int),
        ("bits_per_sample", C.c_int),
        ("source_channels", C.c_int),
        ("use_channel", C.c_int),
        ("extract_pcal", C.c_int),
        ("samplingfreq", C.c_double),
        ("pcaloffsethz", C.c_double),
        ("pcalharmonicshz", C.c_double),
        ("use_channel_file2", C.c_int),
    ]


class FxBuffer(C.Structure):
    """fx_buffer: the reference Buffer (data, allocated bytes, length)."""
    _fields_ = [("data", C.c_void_p), ("allocated", C.c_size_t), ("length", C.c_size_t)]


class SwspecSettings(C.Structure):
    """fx_swspec_settings: one run of the spectrometer application."""
    _fields_ = [
        ("core", SpectrometerSettings),
        ("format", C.c_int),
        ("framebytes", C.c_int),
        ("headerbytes", C.c_int),
        ("frames_per_second", C.c_int),
        ("use_channel_file", C.c_int * 2),
        ("calc_xpol", C.c_int),
        ("spectra_per_chunk", C.c_int),
        ("basefilename", C.c_char_p * 2),
    ]


class SwspecStats(C.Structure):
    _fields_ = [("spectra", C.c_long), ("chunks", C.c_long), ("seconds", C.c_double),
                ("realtime_factor", C.c_double)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    lib = _load_fx()
    vp, i, u, sz, f = C.c_void_p, C.c_int, C.c_uint, C.c_size_t, C.c_float
    TP = C.POINTER(CellDSPTask)
    sigs = {
        "celldsp_start": (i, []),
        "celldsp_stop": (i, []),
        "celldsp_getFreeSPE": (i, []),
        "celldsp_setFreeSPE": (i, [i]),
        "celldsp_startTask": (i, [i, TP]),
        "celldsp_waitTask": (i, [i]),
        "celldsp_executeTask": (i, [TP]),
        "celldsp_startContinuableTask": (i, [TP, C.POINTER(i)]),
        "celldsp_continueTask": (i, [TP, C.POINTER(i)]),
        "celldsp_stopContinuableTask": (i, [C.POINTER(i)]),
        "celldsp_setBehaviour": (None, [u]),
        "celldsp_getBehaviour": (u, []),
        "celldsp_freeCount": (i, []),
        "celldsp_clientCount": (i, []),
        "celldsp_graphReplays": (C.c_long, []),
        "speNOP": (i, []),
        "speIOCheck": (i, []),
        "speIOCheck_out": (i, [vp]),
        "speAddProduct_32fc": (i, [vp, vp, vp, i]),
        "speMul_32fc": (i, [vp, vp, vp, i]),
        "speAdd_32fc": (i, [vp, vp, vp, i]),
        "speFFT_init": (i, [i, C.POINTER(i)]),
        "speFFT_exec": (i, [vp, vp, i, C.POINTER(i)]),
        "speFFT_destroy": (i, [C.POINTER(i)]),
        "speFFT_singleexec": (i, [vp, vp, i]),
        "speDIFX_process": (i, [vp, vp, vp, vp, vp, i, i, i]),
        "speSinCosGenerator": (i, [f, f, vp, i]),
        "speSinCosAccurate": (i, [vp, vp, i]),
        "speSinCosAccurateNorm": (i, [vp, vp, i]),
        "speUnpack2bit": (i, [vp, vp, i, i]),
        "minicorr_run": (i, [C.POINTER(MiniControlBlock), i, C.POINTER(MiniStats)]),
        "spectrometer_cb_run": (i, [C.POINTER(SpecControlBlock), C.POINTER(SpecMbox)]),
        "fx_taskcore_create": (i, [i, C.POINTER(SpectrometerSettings), C.POINTER(vp)]),
        "fx_taskcore_run": (i, [vp, vp, sz, vp, sz, vp, sz]),
        "fx_taskcore_join": (i, [vp]),
        "fx_taskcore_run_buffers": (i, [vp, i, vp, vp, vp, vp]),
        "fx_taskcore_combine_buffers": (i, [vp, vp, vp]),
        "fx_taskcore_combine": (i, [vp, vp, sz, vp, sz]),
        "fx_taskcore_pcal_length": (i, [vp]),
        "fx_taskcore_finalize": (i, [vp]),
        "fx_taskcore_destroy": (i, [vp]),
        "fx_swspectrometer_run": (i, [C.POINTER(SwspecSettings), C.c_char_p, C.c_char_p, C.POINTER(SwspecStats)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    return load().fx_last_error().decode(errors="replace")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _aligned_structs(cls, n=1):
    """n zeroed structures of cls at a 128-byte aligned address (the drop-in's
    control blocks are aligned(128) C types); returns (array, buffer to keep
    alive)."""
    raw = C.create_string_buffer(C.sizeof(cls) * n + 128)
    addr = (C.addressof(raw) + 127) & ~127
    C.memset(addr, 0, C.sizeof(cls) * n)
    return (cls * n).from_address(addr), raw


class CellDSPTaskClass:
    """Task builder (cellspe-tasklib.h:175-222): the struct sits in a
    128-byte aligned slot; vectors are numpy arrays kept alive here."""

    def __init__(self, command):
        self._raw = C.create_string_buffer(C.sizeof(CellDSPTask) + 128)
        addr = (C.addressof(self._raw) + 127) & ~127
        self.task = CellDSPTask.from_address(addr)
        C.memset(addr, 0, C.sizeof(CellDSPTask))
        self.task.command = command
        self.task.completed_flag = CELLDSP_TASK_INCOMPLETE
        self._keep = []

    def addInVec(self, arr):
        t = self.task
        t.inputvector_ptrs[t.num_input_vectors].p = arr.ctypes.data
        t.inputvector_lengths[t.num_input_vectors] = arr.nbytes
        t.num_input_vectors += 1
        self._keep.append(arr)

    def addOutVec(self, arr):
        t = self.task
        t.outputvector_ptrs[t.num_output_vectors].p = arr.ctypes.data
        t.outputvector_lengths[t.num_output_vectors] = arr.nbytes
        t.num_output_vectors += 1
        self._keep.append(arr)

    def setNext(self, other):
        self.task.next_task.p = C.addressof(other.task) if other is not None else None
        self._keep.append(other)

    def setFFTparams(self, nfft):
        self.task.subcontext.FFT.nfft = nfft

    def getTask(self):
        return C.pointer(self.task)

    @property
    def completed_flag(self):
        return self.task.completed_flag

    @property
    def tick_count(self):
        return self.task.tick_count


# ---- minicorrelator ---------------------------------------------------------
def minicorr_run(raw_stations, ffts_to_integrate, fmultipliers=None):
    """Run the control-block drop-in over a list of per-station uint32 word
    arrays (ffts_total*N_FFT/16 words each).  Returns (autos, baselines,
    stats): autos[s] and baselines[b] are (ndump, 2, N_FFT) float32 split
    planes as the PPU reads them."""
    lib = load()
    nst = len(raw_stations)
    nwords = raw_stations[0].size
    T = nwords * 16 // N_FFT
    K = int(ffts_to_integrate)
    ndump = T // K
    nbl = nst * (nst - 1) // 2
    autos = [np.zeros((max(ndump, 1), 2, N_FFT), np.float32) for _ in range(nst)]
    bls = [np.zeros((max(ndump, 1), 2, N_FFT), np.float32) for _ in range(nbl)]
    raws = [np.ascontiguousarray(r, np.uint32) for r in raw_stations]
    fm = None if fmultipliers is None else np.ascontiguousarray(fmultipliers, np.float32)
    cbs, cbs_buf = _aligned_structs(MiniControlBlock, nst)
    for s in range(nst):
        cb = cbs[s]
        cb.rawdata_src.p = raws[s].ctypes.data
        cb.autocorrelation_out.p = autos[s].ctypes.data
        for b in range(nbl):
            cb.baselines_out[b].p = bls[b].ctypes.data
        cb.fmultipliers_src.p = None if fm is None else fm.ctypes.data
        cb.spe_num = s
        cb.num_spes = nst
        cb.ffts_total = T
        cb.ffts_to_integrate = K
    st = MiniStats()
    rc = lib.minicorr_run(cbs, nst, C.byref(st))
    if rc != 0:
        raise FxError(f"minicorr_run: {last_error()}")
    return autos, bls, st


# ---- spectrometer ------------------------------------------------------------
def spectrometer_cb_run(raw_bytes, fft_points, overlap_factor, integrate_ffts, terminate=False):
    """One control-block run; returns (spectra (nspec, N, 2) float32, mbox)."""
    lib = load()
    raw = np.ascontiguousarray(raw_bytes, np.uint8)
    N = int(fft_points)
    ntime = (raw.size // 4) * 16
    hop = N // overlap_factor
    nfft = (ntime - N) // hop + 1 if ntime >= N else 0
    nspec = nfft // integrate_ffts
    dst = np.zeros((max(nspec, 1), N, 2), np.float32)
    cbs, cbs_buf = _aligned_structs(SpecControlBlock)
    cb = cbs[0]
    cb.rank = 0
    cb.terminate_flag = int(terminate)
    cb.fft_points = N
    cb.fft_points_log2 = N.bit_length() - 1
    cb.overlap_factor = overlap_factor
    cb.integrate_ffts = integrate_ffts
    cb.samples_per_rawbyte = 4
    cb.rawif_src.p = raw.ctypes.data
    cb.rawif_bytecount = raw.size
    cb.spectrum_dst.p = dst.ctypes.data
    mb = SpecMbox()
    if lib.spectrometer_cb_run(C.byref(cb), C.byref(mb)) != 0:
        raise FxError(f"spectrometer_cb_run: {last_error()}")
    return dst[: mb.num_spectra], mb


class TaskCoreHIP:
    """TaskCore (TaskCore.h:36-89) over fx_taskcore_*: prepare, asynchronous
    run, join, combineResults, finalize."""

    def __init__(self, rank=0):
        self.rank = rank
        self.handle = None

    def prepare(self, settings: SpectrometerSettings):
        lib = load()
        self.settings = settings
        h = C.c_void_p()
        rc = lib.fx_taskcore_create(self.rank, C.byref(settings), C.byref(h))
        if rc != 0:
            raise FxError(f"TaskCore prepare: {last_error()}")
        self.handle = h
        self.pcal_len = lib.fx_taskcore_pcal_length(h)
        return True

    def run(self, raw, max_spectra):
        s = self.settings
        C1 = s.fft_points // 2 + 1
        nout = 1 if s.use_channel else s.source_channels
        self._raw = np.ascontiguousarray(raw, np.uint8)
        self._spec = np.zeros((max(max_spectra, 1), nout, C1), np.float32)
        self._pcal = np.zeros((max(max_spectra, 1), max(self.pcal_len, 1)), np.complex128)
        rc = load().fx_taskcore_run(self.handle, _p(self._raw), self._raw.nbytes, _p(self._spec),
                                    self._spec.size, _p(self._pcal) if self.pcal_len else None,
                                    self._pcal.shape[0] * self.pcal_len)
        if rc != 0:
            raise FxError(f"TaskCore run: {last_error()}")
        return 0

    def join(self):
        n = load().fx_taskcore_join(self.handle)
        if n < 0:
            raise FxError(f"TaskCore join: {last_error()}")
        self.nspec = n
        return n

    def results(self):
        return self._spec[: self.nspec], self._pcal[: self.nspec, : self.pcal_len]

    @staticmethod
    def _bufs(arrays):
        bufs = [FxBuffer(a.ctypes.data, a.nbytes, 0) for a in arrays]
        ptrs = (C.POINTER(FxBuffer) * len(bufs))(*[C.pointer(b) for b in bufs])
        return bufs, ptrs

    def run_buffers(self, raws, max_spectra, xpol=True, pcal=True):
        """TaskCore::run(inbuf, outbuf, xpolbuf, pcalbuf) over len(raws)
        sources (fx_taskcore_run_buffers); the outputs are numpy arrays sized
        for max_spectra spectra."""
        s = self.settings
        C1 = s.fft_points // 2 + 1
        ns = len(raws)
        npair = ns * (ns - 1) // 2
        m = max(max_spectra, 1)
        self._in = [np.ascontiguousarray(r, np.uint8) for r in raws]
        self._auto = [np.zeros((m, C1), np.float32) for _ in range(ns)]
        self._xpol = [np.zeros((m, C1), np.complex64) for _ in range(npair)] if xpol else []
        self._pcb = [np.zeros((m, max(self.pcal_len, 1)), np.complex64) for _ in range(ns)] \
            if pcal and self.pcal_len else []
        self._ib, ip = self._bufs(self._in)
        for b, a in zip(self._ib, self._in):
            b.length = a.nbytes
        self._ob, op = self._bufs(self._auto)
        self._xb, xp = self._bufs(self._xpol) if self._xpol else ([], None)
        self._pb, pp = self._bufs(self._pcb) if self._pcb else ([], None)
        self._ptrs = (ip, op, xp, pp)
        rc = load().fx_taskcore_run_buffers(self.handle, ns, C.cast(ip, C.c_void_p), C.cast(op, C.c_void_p),
                                            None if xp is None else C.cast(xp, C.c_void_p),
                                            None if pp is None else C.cast(pp, C.c_void_p))
        if rc != 0:
            raise FxError(f"TaskCore run: {last_error()}")
        return 0

    def results_buffers(self):
        """(autos per source, cross spectra per pair, phase-cal per source),
        each cut to the length the core set."""
        C1 = self.settings.fft_points // 2 + 1
        autos = [a[: b.length // (4 * C1)] for a, b in zip(self._auto, self._ob)]
        xp = [a[: b.length // (8 * C1)] for a, b in zip(self._xpol, self._xb)]
        pc = [a[: b.length // (8 * self.pcal_len), : self.pcal_len] for a, b in zip(self._pcb, self._pb)]
        return autos, xp, pc

    def combine_buffers(self, outs, pcals=None):
        """combineResults(outbuf, pcalbuf): outs = autos of each source then
        the cross spectra; returns the lengths the core set."""
        ob, op = self._bufs(outs)
        pb, pp = self._bufs(pcals) if pcals else ([], None)
        rc = load().fx_taskcore_combine_buffers(self.handle, C.cast(op, C.c_void_p),
                                                None if pp is None else C.cast(pp, C.c_void_p))
        if rc != 0:
            raise FxError(f"TaskCore combineResults: {last_error()}")
        return [b.length for b in ob], [b.length for b in pb]

    def combineResults(self, out, pcal=None):
        rc = load().fx_taskcore_combine(self.handle, _p(out), out.size,
                                        None if pcal is None else _p(pcal),
                                        0 if pcal is None else pcal.size)
        if rc != 0:
            raise FxError(f"TaskCore combineResults: {last_error()}")

    def finalize(self):
        return load().fx_taskcore_finalize(self.handle)

    def close(self):
        if self.handle is not None:
            load().fx_taskcore_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

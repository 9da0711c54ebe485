"""GPU parity of the reference-interface seams: the celldsp task runtime and
spe* DSP wrappers (cellspe-tasklib-dsp.h:79-95), the minicorrelator control
block (minicorrelator.h:107-119) and the spectrometer control block and
TaskCore (cell_spectrometer.h:104-115, TaskCore.h:36-89).

Known answers from the reference: IOTEST writes {0.1, 2.3, 4.5, 6.7}
(cellspe-tasklib-dsp.cpp IOTEST handler), the MAC validation of
minitest-cell.cpp:131-161.  Everything else is compared with the CPU oracle
on the same seeded inputs: bit-exact for unpacked samples, normwise relative
error <= 1e-5 for FP32 results.
"""
import ctypes as C
import threading

import numpy as np
import pytest

import oracle as O
import tasklib as TL

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


def nrel(a, b):
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def P(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.fixture(scope="module")
def rt():
    lib = TL.load()
    assert lib.celldsp_start() == 0
    yield lib
    lib.celldsp_stop()


# ---------------------------------------------------------------- spe* wrappers
def test_iocheck(rt, capfd):
    # the reference form: no argument, a static buffer, the values printed
    assert rt.speIOCheck() == 0
    assert "wrote back : vect={0.1,2.3,4.5,6.7}" in capfd.readouterr().out
    out = np.zeros(4, np.float32)
    assert rt.speIOCheck_out(P(out)) == 0
    assert np.array_equal(out, np.array([0.1, 2.3, 4.5, 6.7], np.float32))


def test_mac_validation_kat(rt):
    n = 2048
    i = np.arange(n, dtype=np.float64)
    a = (0.5 * i - 0.5j * i).astype(np.complex64)
    b = (0.15 * i * i - 0.25j).astype(np.complex64)
    acc = np.full(n, 1.0 + 1.5j, np.complex64)
    assert rt.speAddProduct_32fc(P(a), P(b), P(acc), n) == 0
    want = (1 + 0.075 * i**3 - 0.125 * i) + 1j * (1.5 - 0.075 * i**3 - 0.125 * i)
    assert np.max(np.abs(acc - want) / np.maximum(np.abs(want), 1)) < 1e-6
    ref = O.addproduct(a, b, np.full(n, 1.0 + 1.5j, np.complex64), conj=False)
    assert nrel(acc, ref) < 1e-7


def test_mul_add(rt):
    rng = np.random.default_rng(1)
    n = 1000
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    b = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    out = np.zeros(n, np.complex64)
    assert rt.speMul_32fc(P(a), P(b), P(out), n) == 0
    assert nrel(out, O.mul(a, b)) < 1e-7
    assert rt.speAdd_32fc(P(a), P(b), P(out), n) == 0
    assert np.array_equal(out, a + b)


@pytest.mark.parametrize("n", [64, 1024, 16384])
def test_fft_singleexec(rt, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    y = np.zeros(n, np.complex64)
    assert rt.speFFT_singleexec(P(x), P(y), n) == 0
    assert nrel(y, O.fft_r2(x)) < TOL32


def test_fft_continuable(rt):
    n = 4096
    uid = C.c_int(-1)
    free0 = rt.celldsp_freeCount()
    assert rt.speFFT_init(n, C.byref(uid)) == 0
    assert rt.celldsp_freeCount() == free0 - 1
    rng = np.random.default_rng(2)
    for _ in range(3):
        x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        y = np.zeros(n, np.complex64)
        assert rt.speFFT_exec(P(x), P(y), n, C.byref(uid)) == 0
        assert nrel(y, O.fft_r2(x)) < TOL32
    assert rt.speFFT_destroy(C.byref(uid)) == 0
    assert rt.celldsp_freeCount() == free0


def test_sincos(rt):
    rng = np.random.default_rng(3)
    n = 5000
    arg = rng.uniform(-50, 50, n).astype(np.float32)
    out = np.zeros(n, np.complex64)
    assert rt.speSinCosAccurate(P(arg), P(out), n) == 0
    want = np.exp(1j * arg.astype(np.float64))
    assert np.max(np.abs(out - want)) < 4e-6
    turns = rng.uniform(-3, 3, n).astype(np.float32)
    assert rt.speSinCosAccurateNorm(P(turns), P(out), n) == 0
    want = np.exp(2j * np.pi * turns.astype(np.float64))
    assert np.max(np.abs(out - want)) < 4e-6


def test_sincos_generator(rt):
    n = 100000
    start, inc = np.float32(0.3), np.float32(0.0123)
    out = np.zeros(n, np.complex64)
    assert rt.speSinCosGenerator(C.c_float(start), C.c_float(inc), P(out), n) == 0
    ph = np.float64(start) + np.arange(n) * np.float64(inc)
    assert np.max(np.abs(out - np.exp(1j * ph))) < 2e-6


@pytest.mark.parametrize("nchan", [1, 16])
def test_unpack_wrapper_bitexact(rt, nchan):
    rng = np.random.default_rng(4 + nchan)
    nsamp = 4096
    words = rng.integers(0, 2**32, nsamp * nchan // 16, dtype=np.uint64).astype(np.uint32)
    out = np.zeros(nsamp * nchan, np.float32)
    assert rt.speUnpack2bit(P(words), P(out), nsamp, nchan) == 0
    ref = O.unpack_1ch(words, nsamp) if nchan == 1 else O.unpack_nch(words, nsamp, nchan).ravel()
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("upper", [0, 1])
def test_difx_process(rt, upper):
    nfft, nch = 1024, 512
    rng = np.random.default_rng(7 + upper)
    x = (rng.standard_normal(nfft) + 1j * rng.standard_normal(nfft)).astype(np.complex64)
    nproc = nch + 1 + 3
    mult = np.exp(1j * rng.uniform(0, 2 * np.pi, nproc)).astype(np.complex64)
    fo = np.zeros(nproc, np.complex64)
    cfo = np.zeros(nproc, np.complex64)
    acc0 = (rng.standard_normal(nproc)).astype(np.float32) + 0j
    acc = acc0.astype(np.complex64)
    assert rt.speDIFX_process(P(x), P(mult), P(fo), P(cfo), P(acc), nfft, 2 * nch, upper) == 0
    X = O.fft_r2(x).astype(np.complex128)
    band0 = nch if upper else 0
    k = np.arange(nch + 1)
    y = X[(band0 + k) % nfft] * mult[: nch + 1]
    assert nrel(fo[: nch + 1], y) < TOL32
    assert np.array_equal(cfo[: nch + 1], np.conj(fo[: nch + 1]))
    assert nrel(acc[: nch + 1].real, acc0[: nch + 1].real + np.abs(y) ** 2) < TOL32
    assert np.array_equal(acc[: nch + 1].imag, acc0[: nch + 1].imag.astype(np.float32))


def _graph_chain():
    """FFT -> MULC -> MAC, an UNPACK, and a DIFX task fed by the FFT output
    through its host vector: five tasks, data passed between them on the host."""
    n, nch = 1024, 256
    rng = np.random.default_rng(31)
    v = {
        "x": np.zeros(n, np.complex64), "X": np.zeros(n, np.complex64),
        "m": np.exp(1j * rng.uniform(0, 2 * np.pi, n)).astype(np.complex64),
        "Y": np.zeros(n, np.complex64), "acc": np.zeros(n, np.complex64),
        "words": np.zeros(512, np.uint32), "f": np.zeros(4096 * 2, np.float32),
        "mult": np.exp(1j * rng.uniform(0, 2 * np.pi, nch + 1)).astype(np.complex64),
        "fo": np.zeros(nch + 1, np.complex64), "cfo": np.zeros(nch + 1, np.complex64),
        "acc2": np.zeros(nch + 1, np.complex64),
    }
    t1 = TL.CellDSPTaskClass(TL.CELLDSP_TASK_FFT_EXEC_DYN)
    t1.setFFTparams(n)
    t1.addInVec(v["x"]); t1.addOutVec(v["X"])
    t2 = TL.CellDSPTaskClass(TL.CELLDSP_TASK_MULC)
    t2.addInVec(v["X"]); t2.addInVec(v["m"]); t2.addOutVec(v["Y"])
    t3 = TL.CellDSPTaskClass(TL.CELLDSP_TASK_MAC)
    t3.addInVec(v["Y"]); t3.addInVec(v["m"]); t3.addOutVec(v["acc"])
    t4 = TL.CellDSPTaskClass(TL.CELLDSP_TASK_UNPACK)
    t4.task.subcontext.Unpack.nchan = 2
    t4.addInVec(v["words"]); t4.addOutVec(v["f"])
    t5 = TL.CellDSPTaskClass(TL.CELLDSP_TASK_DIFXPROCESS)
    d = t5.task.subcontext.Difx
    d.nfft, d.numchannels, d.useUpperOrLower = n, nch, 1
    t5.addInVec(v["X"]); t5.addInVec(v["mult"])
    t5.addOutVec(v["fo"]); t5.addOutVec(v["cfo"]); t5.addOutVec(v["acc2"])
    tasks = [t1, t2, t3, t4, t5]
    for a, b in zip(tasks, tasks[1:]):
        a.setNext(b)
    return tasks, v


def test_task_chain_graph_replays(rt):
    """CDSP_BEHAV_GRAPH: the repeated chain runs from its captured graph and
    every replay equals the eager execution bit for bit (inputs changed in
    place between replays, eager runs and synchronizations in between)."""
    tasks, v = _graph_chain()
    rng = np.random.default_rng(32)
    outs = ("X", "Y", "acc", "f", "fo", "cfo", "acc2")
    r0 = rt.celldsp_graphReplays()
    for rep in range(5):
        x = (rng.standard_normal(1024) + 1j * rng.standard_normal(1024)).astype(np.complex64)
        words = rng.integers(0, 2**32, 512, dtype=np.uint64).astype(np.uint32)
        acc0 = (rng.standard_normal(1024) + 0j).astype(np.complex64)
        acc20 = (rng.standard_normal(257) + 0j).astype(np.complex64)
        got = {}
        for behaviour in (TL.CDSP_BEHAV_GRAPH, TL.CDSP_BEHAV_DEFAULT):
            v["x"][:] = x; v["words"][:] = words; v["acc"][:] = acc0; v["acc2"][:] = acc20
            for k in ("X", "Y", "f", "fo", "cfo"):
                v[k][:] = 0
            rt.celldsp_setBehaviour(behaviour)
            assert rt.celldsp_executeTask(tasks[0].getTask()) == 0
            assert all(t.completed_flag == TL.CELLDSP_TASK_OK for t in tasks)
            got[behaviour] = {k: v[k].copy() for k in outs}
            if behaviour == TL.CDSP_BEHAV_GRAPH and rep > 0:
                # replayed: the chain's device time is shared evenly
                assert len({t.tick_count for t in tasks}) == 1
        for k in outs:
            assert np.array_equal(got[TL.CDSP_BEHAV_GRAPH][k].view(np.uint32),
                                  got[TL.CDSP_BEHAV_DEFAULT][k].view(np.uint32)), (rep, k)
        assert nrel(got[TL.CDSP_BEHAV_DEFAULT]["X"], O.fft_r2(x)) < TOL32
    # the first graph-mode execution runs eagerly and sizes the buffers; the
    # next four replay
    assert rt.celldsp_graphReplays() - r0 == 4
    # a changed chain (another multiplier vector) runs eagerly once, then replays
    other = v["m"].copy()
    tasks[1].task.inputvector_ptrs[1].p = other.ctypes.data
    rt.celldsp_setBehaviour(TL.CDSP_BEHAV_GRAPH)
    for k in range(2):
        assert rt.celldsp_executeTask(tasks[0].getTask()) == 0
    assert rt.celldsp_graphReplays() - r0 == 5
    rt.celldsp_setBehaviour(TL.CDSP_BEHAV_DEFAULT)


def test_chain_graph_after_slot_growth(rt):
    """A captured chain's graph holds the unit's device staging addresses.  A
    larger single task on the same unit regrows one of those buffers; the
    chain's next graph-mode execution must not replay the stale graph: it
    runs eagerly, is captured again on the one after, and every result equals
    the eager chain's bit for bit."""
    tasks, v = _graph_chain()
    rng = np.random.default_rng(34)
    outs = ("X", "Y", "acc", "f", "fo", "cfo", "acc2")
    unit = rt.celldsp_getFreeSPE()
    assert unit >= 0

    def run(head):
        assert rt.celldsp_startTask(unit, head) == 0
        return rt.celldsp_waitTask(unit)

    def chain(behaviour, x, words, acc0, acc20):
        v["x"][:] = x; v["words"][:] = words; v["acc"][:] = acc0; v["acc2"][:] = acc20
        for k in ("X", "Y", "f", "fo", "cfo"):
            v[k][:] = 0
        rt.celldsp_setBehaviour(behaviour)
        assert run(tasks[0].getTask()) == TL.CELLDSP_TASK_OK
        return {k: v[k].copy() for k in outs}

    def inputs():
        return ((rng.standard_normal(1024) + 1j * rng.standard_normal(1024)).astype(np.complex64),
                rng.integers(0, 2**32, 512, dtype=np.uint64).astype(np.uint32),
                (rng.standard_normal(1024) + 0j).astype(np.complex64),
                (rng.standard_normal(257) + 0j).astype(np.complex64))

    try:
        for _ in range(3):              # eager (sizes the buffers), capture + replay, replay
            chain(TL.CDSP_BEHAV_GRAPH, *inputs())
        r0 = rt.celldsp_graphReplays()
        # a 32 MiB output round trip through HBM: larger than anything the
        # tests before put in that output buffer, so it is reallocated
        big = TL.CellDSPTaskClass(TL.CELLDSP_TASK_IOBENCH)
        vb = rng.integers(0, 2**32, 8 << 20, dtype=np.uint64).astype(np.uint32)
        sent = vb.copy()
        big.addOutVec(vb)
        assert run(big.getTask()) == TL.CELLDSP_TASK_OK
        assert np.array_equal(vb, sent)
        for rep in range(3):
            args = inputs()
            got = chain(TL.CDSP_BEHAV_GRAPH, *args)
            ref = chain(TL.CDSP_BEHAV_DEFAULT, *args)
            for k in outs:
                assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), (rep, k)
            assert nrel(ref["X"], O.fft_r2(args[0])) < TOL32
        # after the growth: one eager run, then a new capture and two replays
        assert rt.celldsp_graphReplays() - r0 == 2
    finally:
        rt.celldsp_setBehaviour(TL.CDSP_BEHAV_DEFAULT)
        rt.celldsp_setFreeSPE(unit)


def test_task_chain_on_device(rt):
    io_out = np.zeros(4, np.float32)
    t1 = TL.CellDSPTaskClass(TL.CELLDSP_TASK_IOTEST)
    t1.addOutVec(io_out)
    n = 256
    a = np.ones(n, np.complex64) * (1 + 2j)
    b = np.ones(n, np.complex64) * (3 - 1j)
    acc = np.zeros(n, np.complex64)
    t2 = TL.CellDSPTaskClass(TL.CELLDSP_TASK_MAC)
    t2.addInVec(a)
    t2.addInVec(b)
    t2.addOutVec(acc)
    t1.setNext(t2)
    assert rt.celldsp_executeTask(t1.getTask()) == 0
    assert t1.completed_flag == 0 and t2.completed_flag == 0
    assert np.array_equal(io_out, np.array([0.1, 2.3, 4.5, 6.7], np.float32))
    assert np.allclose(acc, (1 + 2j) * (3 - 1j))
    assert t1.tick_count >= 0


def test_parallel_clients(rt):
    # several host threads sharing the units, as TaskDispatcher workers would
    errs = []

    def work(seed):
        rng = np.random.default_rng(seed)
        for _ in range(5):
            x = (rng.standard_normal(512) + 1j * rng.standard_normal(512)).astype(np.complex64)
            y = np.zeros(512, np.complex64)
            if rt.speFFT_singleexec(P(x), P(y), 512) != 0 or nrel(y, O.fft_r2(x)) > TOL32:
                errs.append(seed)

    rt.celldsp_setBehaviour(TL.CDSP_BEHAV_BLOCKING)
    th = [threading.Thread(target=work, args=(s,)) for s in range(12)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    rt.celldsp_setBehaviour(TL.CDSP_BEHAV_DEFAULT)
    assert not errs
    assert rt.celldsp_freeCount() == TL.CELLDSP_MAX_SPE_NR


# ---------------------------------------------------------------- minicorrelator
@pytest.mark.parametrize("nst,with_mult", [(4, False), (6, True)])
def test_minicorr_control_block(nst, with_mult):
    N = TL.N_FFT
    T, K = 96, 32
    st = [O.gen_station(77, s, 0.3, s, 0.0, 0.0, 0.9816, T * N) for s in range(nst)]
    fm = np.linspace(0, 3.7, N // 2).astype(np.float32) if with_mult else None
    autos, bls, stats = TL.minicorr_run(st, K, fm)
    assert stats.dumps == T // K and stats.spectra == T
    C_ = N // 2
    wpb = K * N // 16
    bl = 0
    for b in range(T // K):
        ref = O.corr_f32([s[b * wpb:(b + 1) * wpb] for s in st], N, K)
        p = 0
        for i in range(nst):
            for j in range(i, nst):
                buf = autos[i][b] if i == j else bls[i * nst - i * (i + 1) // 2 + (j - i - 1)][b]
                v = buf[0] + 1j * buf[1]
                assert nrel(v[:C_], ref[p]) < TOL32, (b, i, j)
                assert v[C_] == 0
                assert nrel(v[C_ + 1:], np.conj(v[1:C_][::-1])) == 0
                p += 1
    del bl


# ---------------------------------------------------------------- spectrometer
@pytest.mark.parametrize("N,ov,K", [(1024, 2, 8), (4096, 1, 3), (32768, 2, 2)])
def test_spectrometer_control_block(N, ov, K):
    ntime = N * 40
    w = O.gen_multichannel(5, 1, [0.2], [0.123], 0.9816, ntime)
    spec, mb = TL.spectrometer_cb_run(w.view(np.uint8), N, ov, K)
    ref = O.spectrometer_ex_f32(w, ntime, 1, N, ov, K, O.WIN_SIN2, True, True, nthreads=8)
    assert mb.num_spectra == ref.shape[0] and mb.total_ffts == (ntime - N) // (N // ov) + 1
    assert np.all(spec[:, :, 1] == 0)
    p = spec[:, :, 0]
    assert nrel(p[:, : N // 2 + 1], ref[:, 0, :]) < TOL32
    assert np.array_equal(p[:, N // 2 + 1:], p[:, 1: N // 2][:, ::-1])


def _settings(N, ov, K, win, nch, use, pcal=0, fs=16e6, off=500e3, harm=1e6, use2=0):
    return TL.SpectrometerSettings(fft_points=N, fft_overlap_factor=ov, core_overlapped_ffts=K,
                                   window=win, bits_per_sample=2, source_channels=nch,
                                   use_channel=use, extract_pcal=pcal, samplingfreq=fs,
                                   pcaloffsethz=off, pcalharmonicshz=harm, use_channel_file2=use2)


@pytest.mark.parametrize("nch,use,win", [(1, 0, O.WIN_HANN), (4, 0, O.WIN_COSINE), (4, 3, O.WIN_BLACKMAN)])
def test_taskcore_spectra(nch, use, win):
    N, ov, K = 2048, 2, 6
    ntime = N * 30 + 123 * 16 // nch
    amps = [0.1 * (c + 1) for c in range(nch)]
    freqs = [0.05 + 0.07 * c for c in range(nch)]
    w = O.gen_multichannel(9, nch, amps, freqs, 0.9816, ntime)
    core = TL.TaskCoreHIP(rank=0)
    core.prepare(_settings(N, ov, K, win, nch, use))
    core.run(w.view(np.uint8), max_spectra=64)
    n = core.join()
    ref = O.spectrometer_ex_f32(w, ntime, nch, N, ov, K, win, False, True, nthreads=8)
    assert n == ref.shape[0]
    spec, _ = core.results()
    want = ref[:, use - 1: use, :] if use else ref
    assert nrel(spec, want) < TOL32
    acc = np.ones_like(spec)
    core.combineResults(acc)
    assert np.array_equal(acc, spec + 1)
    core.finalize()
    core.close()


def test_taskcore_pcal_and_two_cores():
    N, ov, K = 1024, 2, 5
    fs, off, harm = 16e6, 500e3, 1e6
    T, R = 16, 32
    half = N * 24
    amps, freqs = [0.3], [(off + 3 * harm) / fs]
    w = O.gen_multichannel(13, 1, amps, freqs, 0.9816, 2 * half)
    cores = [TL.TaskCoreHIP(rank=r) for r in range(2)]
    for c in cores:
        c.prepare(_settings(N, ov, K, O.WIN_SIN2, 1, 0, pcal=1, fs=fs, off=off, harm=harm))
        assert c.pcal_len == T
    # TaskDispatcher style: each core gets its own slice, runs concurrently
    raw = w.view(np.uint8)
    slices = [raw[: half // 4], raw[half // 4:]]
    for c, s in zip(cores, slices):
        c.run(s, max_spectra=64)
    counts = [c.join() for c in cores]
    x = O.unpack_1ch(w, 2 * half)
    L = (K - 1) * (N // ov) + N
    hop = N // ov
    for c, n, base in zip(cores, counts, (0, half)):
        spec, pc = c.results()
        assert n == half // L
        ref = O.spectrometer_ex_f32(w[base // 16:(base + half) // 16], half, 1, N, ov, K, O.WIN_SIN2,
                                    False, True)
        assert nrel(spec, ref) < TOL32
        for s in range(n):
            acc = np.zeros(T, np.complex128)
            for j in range(K):
                if (s * K + j) % ov == 0:
                    st = base + s * L + j * hop
                    O.tcpcal_fold(x[st: st + N], T, R, -2 * np.pi * off / fs, acc)
            assert nrel(pc[s], acc) < 1e-12
    total = np.zeros_like(cores[0].results()[0])
    for c in cores:
        c.combineResults(total)
    assert nrel(total, cores[0].results()[0] + cores[1].results()[0]) < 1e-7
    for c in cores:
        c.close()


def _power_sum_f64(x, N, hop, nfft, win):
    """FP64 sum over nfft windowed FFTs of x (hop apart) of |X|^2, bins 0..N/2."""
    w = O.window(win, N).astype(np.float64)
    return sum(np.abs(np.fft.rfft(x[j * hop: j * hop + N] * w)) ** 2 for j in range(nfft))


def _cross_sum_f64(xa, xb, N, hop, nfft, win):
    """FP64 sum of X_a conj(X_b) with DC and Nyquist the real-part products."""
    w = O.window(win, N).astype(np.float64)
    acc = np.zeros(N // 2 + 1, np.complex128)
    for j in range(nfft):
        A = np.fft.rfft(xa[j * hop: j * hop + N] * w)
        B = np.fft.rfft(xb[j * hop: j * hop + N] * w)
        p = A * np.conj(B)
        p[0] = A[0].real * B[0].real
        p[-1] = A[-1].real * B[-1].real
        acc += p
    return acc


def test_taskcore_buffers_two_sources_xpol_partial():
    """TaskCore::run(inbuf, outbuf, xpolbuf, pcalbuf) with two 4-channel
    sources (channel 2 of the first, channel 3 of the second): the mean
    power of every full spectrum, the cross spectrum of the pair, and at the
    shorter source's end a partial spectrum of 3 FFTs holding unscaled sums
    (TaskCoreIPP.cpp:456-487,519-524); combineResults adds each source's and
    the pair's first spectrum."""
    N, ov, K, nch, win = 1024, 2, 5, 4, O.WIN_HANN
    hop = N // ov
    L = (K - 1) * hop + N
    f = 3
    ntime0 = 3 * L + (f - 1) * hop + N + 48          # 3 full spectra, then 3 whole FFTs and a bit
    ntime1 = ntime0 + 5000
    w0 = O.gen_multichannel(41, nch, [0.2, 0.1, 0.0, 0.3], [0.11, 0.2, 0.0, 0.31], 0.9816, ntime0)
    w1 = O.gen_multichannel(42, nch, [0.0, 0.1, 0.2, 0.3], [0.0, 0.2, 0.11, 0.31], 0.9816, ntime1)
    core = TL.TaskCoreHIP(rank=0)
    core.prepare(_settings(N, ov, K, win, nch, 2, use2=3))
    core.run_buffers([w0.view(np.uint8), w1.view(np.uint8)], max_spectra=8)
    n = core.join()
    assert n == 4
    autos, xps, _ = core.results_buffers()
    assert [a.shape[0] for a in autos] == [4, 4] and len(xps) == 1 and xps[0].shape[0] == 4
    ch = (1, 2)
    for k, w in enumerate((w0, w1)):
        ref = O.spectrometer_ex_f32(w, 3 * L, nch, N, ov, K, win, False, True)[:, ch[k]]
        assert nrel(autos[k][:3], ref) < TOL32
        x = O.unpack_nch(w, ntime0, nch)[ch[k]].astype(np.float64)[3 * L:]
        assert nrel(autos[k][3], _power_sum_f64(x, N, hop, f, win)) < TOL32
    xref = O.spectrometer_xpol_f64(w0, nch, ch[0], w1, nch, ch[1], 3 * L, N, ov, K, win)
    assert nrel(xps[0][:3], xref) < TOL32
    xa = O.unpack_nch(w0, ntime0, nch)[ch[0]].astype(np.float64)[3 * L:]
    xb = O.unpack_nch(w1, ntime0, nch)[ch[1]].astype(np.float64)[3 * L:]
    assert nrel(xps[0][3], _cross_sum_f64(xa, xb, N, hop, f, win)) < TOL32
    outs = [np.zeros(N // 2 + 1, np.float32), np.full(N // 2 + 1, 1.0, np.float32), np.zeros(N // 2 + 1, np.complex64)]
    lens, _ = core.combine_buffers(outs)
    assert lens == [4 * (N // 2 + 1), 4 * (N // 2 + 1), 8 * (N // 2 + 1)]
    assert np.array_equal(outs[0], autos[0][0]) and np.array_equal(outs[1], autos[1][0] + 1)
    assert np.array_equal(outs[2], xps[0][0])
    core.finalize()
    core.close()


def test_taskcore_buffers_pcal_partial():
    """Phase-cal per source through the Buffer form: complex float32 tone bins
    per spectrum (FP64 accumulation), the partial spectrum folding only its
    non-overlapped FFTs; no cross spectrum requested."""
    N, ov, K = 1024, 2, 5
    fs, off, harm = 16e6, 500e3, 1e6
    T, R = 16, 32
    hop = N // ov
    L = (K - 1) * hop + N
    f = 4
    ntime = 2 * L + (f - 1) * hop + N
    ws = [O.gen_multichannel(50 + k, 1, [0.3], [(off + (3 + k) * harm) / fs], 0.9816, ntime) for k in range(2)]
    core = TL.TaskCoreHIP(rank=1)
    core.prepare(_settings(N, ov, K, O.WIN_SIN2, 1, 0, pcal=1, fs=fs, off=off, harm=harm))
    core.run_buffers([w.view(np.uint8) for w in ws], max_spectra=4, xpol=False)
    assert core.join() == 3
    autos, xps, pcs = core.results_buffers()
    assert xps == [] and [p.shape for p in pcs] == [(3, T), (3, T)]
    for k, w in enumerate(ws):
        x = O.unpack_1ch(w, ntime)
        for s in range(3):
            acc = np.zeros(T, np.complex128)
            for j in range(K if s < 2 else f):
                if (s * K + j) % ov == 0:
                    st = s * L + j * hop
                    O.tcpcal_fold(x[st: st + N], T, R, -2 * np.pi * off / fs, acc)
            assert nrel(pcs[k][s], acc) < 1e-6
        ref = O.spectrometer_ex_f32(w, 2 * L, 1, N, ov, K, O.WIN_SIN2, False, True)[:, 0]
        assert nrel(autos[k][:2], ref) < TOL32
    core.close()

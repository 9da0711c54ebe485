#!/usr/bin/env python3
"""Benchmark of the F-X chain on one MI355X.

Headline workload (the mid-size correlator): 64 stations x 2 polarizations =
128 inputs of 2-bit samples with correlated fraction rho = 0.02 and seeded
per-station delays of 0-50 samples, 8192-point real FFT (4096 channels), FP64
delay/fringe phase model applied after the FFT, 65536 spectra (2^29 samples
per input, 16 GiB of packed input resident in HBM), all 8256 products.

One step = reset accumulators -> per 16384-spectrum chunk: F kernel (unpack,
FFT, rotate, scale, convert, store in the engine layout) -> X-engine -> dump
[p][c] complex64 visibilities.  HIP events around the launches of the first
chunk give the per-kernel durations inside the timed region.  The headline
X-engine is FP16 MFMA; the plain FP32 path and the scaled FP8 path run the
same step beside it.

Further sections: 2-bit unpack, the 16-channel spectrometer workload (2^28
samples, 32768-point FFTs), the visibility products (lag spectra, amplitude/phase),
the VDIF front end (frame placement of ~1 GB of frames in HBM, and the
double-buffered file ingest), a task-runtime chain per task vs as one
hipGraph, the spectrometer application over two files, and the batch of
uneven subbands.

A second section times the large-N X-engine (256 stations x 2 pol = 512
inputs, 1024 channels, 8192 spectra of 2-bit channelized voltages): the exact
FP8 decomposition engine against FP16 and plain FP32 MFMA engines.

Prints ONE JSON line.  `value` is whole-step complex MAC/s (P*C*T / step time).
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cellspe-tasklib_amd"))

METRIC = ("X-engine complex MAC/s + input GS/s, FP16/FP8 MFMA vs FP32 plain at ≤1e-3 err")
PEAK = {"hbm": 8000.0, "fp16": 2500.0, "fp32": 157.3, "fp8": 5000.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(args):
    I = 2 * args.stations
    N = args.nfft
    C = N // 2
    T = args.spectra
    rng = np.random.default_rng(args.seed)
    tau = rng.uniform(0.0, 50.0, args.stations)
    fr0 = rng.uniform(-np.pi, np.pi, args.stations)
    fr1 = rng.uniform(-1e-3, 1e-3, args.stations)
    models = []
    for s in range(args.stations):
        for _pol in range(2):
            models.append((float(tau[s]), 1e-7 * s, float(fr0[s]), float(fr1[s])))
    return I, N, C, T, tau, models


def nrel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def cpu_threads():
    """Threads of the CPU baseline: the CPUs this process may run on, at most
    16 (one GPU's share of the host on the benchmark boxes).  Returns
    (threads, CPUs in the affinity mask, online CPUs)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    online = os.sysconf("SC_NPROCESSORS_ONLN")
    return min(16, aff), aff, online


def cpu_median(fn, reps=3):
    """Median wall time of reps calls of fn after one warm-up call."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_obj(value, unit, sample, extra=None):
    nth, aff, online = cpu_threads()
    o = {"value": value, "unit": unit, "cores": nth, "kind": "port", "sample": sample,
         "affinity_cpus": aff, "online_cpus": online,
         "build": "oracle/fx_oracle.c -O3 -march=native -fopenmp on this host; project radix-2 FFT, not FFTW",
         "timing": "median of 3 runs after one warm-up"}
    if extra:
        o.update(extra)
    return o


def measured_busy(kernel, unit="valu"):
    """VALU (or MFMA) busy fraction of `kernel` from the committed rocprofv3
    counter summaries (lines '<unit>_busy <kernel> <fraction>'), or None."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for fn in sorted(os.listdir(pdir)):
        with open(os.path.join(pdir, fn)) as f:
            for line in f:
                m = re.match(r"^" + unit + r"_busy\s+(\S.*\S)\s+([0-9.eE+]+)\s*$", line)
                if m and m.group(1) == kernel:
                    return {"frac": float(m.group(2)), "source": f"profiles/{fn}"}
    return None


def measured_valu(kernel):
    return measured_busy(kernel, "valu")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summaries (lines 'bytes_per_launch <kernel> <bytes>'), or None."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for fn in sorted(os.listdir(pdir)):
        if "traffic" not in fn:
            continue
        with open(os.path.join(pdir, fn)) as f:
            for line in f:
                m = re.match(r"^bytes_per_launch\s+(\S.*\S)\s+([0-9.eE+]+)\s*$", line)
                if m and m.group(1) == kernel:
                    return {"bytes": float(m.group(2)), "source": f"profiles/{fn}"}
    return None


def copy_probe(fxlib, args):
    ctx = fxlib.Context(0)
    nbytes = args.copy_probe << 20
    src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
    src.zero()
    lib = fxlib.load()
    for _ in range(args.warmup):
        fxlib.check(lib.fx_copy_kernel(ctx.handle, src.ptr, dst.ptr, nbytes))
    e0, e1 = ctx.event(), ctx.event()
    ctx.sync()
    ctx.record(e0)
    for _ in range(args.steps):
        fxlib.check(lib.fx_copy_kernel(ctx.handle, src.ptr, dst.ptr, nbytes))
    ctx.record(e1)
    ms = ctx.elapsed_ms(e0, e1) / args.steps
    print(json.dumps({"copy_probe_bytes": nbytes, "ms_per_copy": ms,
                      "gbps_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
    ctx.close()


def mini_correlator(fxlib, ctx, args, O):
    """The plain mini FX correlator (FP32): 4 stations, 2^20 samples each,
    0.05 correlated fraction with seeded integer delays 0-7, a 10 kHz
    phase-cal tone at 0.03 sigma (fs = 16 MHz), 1024-point real FFT (512
    channels), all 10 products over all 1024 spectra; inputs resident in HBM.
    GPU: reset + F + X + dump per step (fx_corr_process, FP32 MFMA engine);
    CPU: the FP32 restatement of the same chain on the same words."""
    nst, nsamp, n = 4, 1 << 20, 1024
    T, C = nsamp // n, n // 2
    P = nst * (nst + 1) // 2
    delays = np.random.default_rng(args.seed + 1).integers(0, 8, nst)
    nw = nsamp // 16
    dw = ctx.alloc(nst * nw * 4)
    lib = fxlib.load()
    for s in range(nst):
        fxlib.check(lib.fx_synth_station(ctx.handle, args.seed + 1, s, 0.05, int(delays[s]), 0.03, 10e3 / 16e6,
                                         0.9816, nsamp, dw.ptr.value + s * nw * 4))
    corr = fxlib.Correlator(ctx, nst, n, T, fxlib.FX_PREC_F32)
    dvis = ctx.alloc(P * C * 8)

    def step():
        corr.reset()
        corr.process(dw, nw, 0, T)
        corr.dump(dvis)

    step()
    ctx.sync()
    reps = max(20, args.steps)
    e = [ctx.event() for _ in range(2)]
    ctx.record(e[0])
    for _ in range(reps):
        step()
    ctx.record(e[1])
    ctx.sync()
    ms = ctx.elapsed_ms(e[0], e[1]) / reps
    cmac = float(P) * C * T
    out = {"workload": "4 stations x 2^20 2-bit samples, rho=0.05, delays 0-7, 10 kHz pcal tone, 1024-pt FFT, "
                       "512 channels, 10 products over 1024 spectra, FP32",
           "ms": ms, "complex_mac_per_s": cmac / (ms * 1e-3), "input_gsps": nst * nsamp / (ms * 1e-3) / 1e9,
           "xengine_kernel": corr.xengine_name(T)}
    got = dvis.download(np.complex64, shape=(P, C))
    corr.close()
    dvis.free()
    if not args.no_cpu:
        nth = cpu_threads()[0]
        streams = [dw.download_at(s * nw * 4, np.uint32, nw) for s in range(nst)]
        ref = O.corr_f32(streams, n, T, nthreads=nth)
        out["err_vs_cpu_fp32"] = nrel(got, ref)
        dt = cpu_median(lambda: O.corr_f32(streams, n, T, nthreads=nth))
        out["cpu_baseline"] = cpu_obj(cmac / dt, "complex MAC/s", "the whole workload", {
            "ms": dt * 1e3, "input_gsps": nst * nsamp / dt / 1e9})
        out["gpu_over_cpu"] = dt * 1e3 / ms
    log(f"[bench] mini correlator: {ms:.3f} ms ({out['input_gsps']:.1f} GS/s)" +
        (f", CPU {out['cpu_baseline']['ms']:.1f} ms" if "cpu_baseline" in out else ""))
    dw.free()
    return out


def large_n(fxlib, ctx, args):
    """Large-N X-engine: 2-bit channelized voltages, exact FP8 decomposition
    vs FP16 and plain FP32 engines (X-engine time only)."""
    I, C, T = args.large_inputs, args.large_channels, args.large_spectra
    P = I * (I + 1) // 2
    cmac = float(P) * C * T
    log(f"[bench] large-N X-engine: I={I} C={C} T={T} P={P}")
    codes = ctx.alloc(C * I * T)
    ctx.synth_channelized(args.seed + 3, C, T, I, 0.01, 0.9816, codes, input_major=True)
    vis = ctx.alloc(P * C * 8)
    out = {"workload": f"{I} inputs (256 stations x 2 pol), {C} channels, {T} spectra of 2-bit "
                       f"channelized voltages, rho=0.01", "products": P}
    sample_ch = [0, C // 3, 2 * C // 3, C - 1]
    vis_sample = {}

    def timed(corr, dx, reps):
        e0, e1 = ctx.event(), ctx.event()
        corr.reset()
        corr.xengine(dx, T)          # warm-up
        ctx.sync()
        corr.reset()
        ctx.record(e0)
        for _ in range(reps):
            corr.xengine(dx, T)
        ctx.record(e1)
        ctx.sync()
        ms = ctx.elapsed_ms(e0, e1) / reps
        corr.reset()
        corr.xengine(dx, T)
        corr.dump(vis)
        ctx.sync()
        v = np.stack([vis.download_at(0, np.complex64, P * C).reshape(P, C)[:, c] for c in sample_ch], 1)
        return ms, v

    res = {}
    for name, prec in (("fp4_2bit", fxlib.FX_PREC_FP4_2BIT), ("fp8_2bit", fxlib.FX_PREC_FP8_2BIT)):
        c8 = fxlib.Correlator(ctx, I, 2 * C, T, prec)
        ms, v = timed(c8, codes, args.steps)
        c8.close()
        res[name], vis_sample[name] = ms, v
        log(f"[bench] large-N {name} (exact 2-bit decomposition): {ms:.2f} ms")
    knames = {"fp4_2bit": "k_xengine4 (FP4 2-bit)", "fp8_2bit": "k_xengine8<3> (FP8 2-bit)"}
    # f16_lockstep: the FP16 engine in lockstep groups (HBM fetches of the
    # shared input blocks once instead of about three times; bit-identical)
    for name, prec, esz, reps, kind in (("f16", fxlib.FX_PREC_F16, 4, args.steps, fxlib.FX_XENGINE_AUTO),
                                        ("f16_lockstep", fxlib.FX_PREC_F16, 4, args.steps,
                                         fxlib.FX_XENGINE_LOCKSTEP),
                                        ("f32", fxlib.FX_PREC_F32, 8, 1, fxlib.FX_XENGINE_AUTO)):
        if name == "f32" and args.no_fp32:
            continue
        dx = ctx.alloc(C * T * I * esz)
        ctx.codes_to_chan(codes, True, C, T, I, prec, dx)
        cc = fxlib.Correlator(ctx, I, 2 * C, 8, prec)
        # the level values go in unscaled: unit power scales in the dump
        cc.set_scales(np.ones(C, np.float32), np.ones(I, np.float32))
        cc.set_xengine(kind)
        knames[name] = cc.xengine_name(T, dx)
        ms, v = timed(cc, dx, reps)
        cc.close()
        dx.free()
        res[name], vis_sample[name] = ms, v
        log(f"[bench] large-N {name}: {ms:.2f} ms ({knames[name]})")
    # FP64 reference on the sample channels (numpy, level-table values)
    lut = np.array([-3.3359, -1.0, 1.0, 3.3359], np.float32).astype(np.float64)
    iu = np.triu_indices(I)
    err = {}
    if not args.no_check:
        ref = np.zeros((P, len(sample_ch)), np.complex128)
        for k, c in enumerate(sample_ch):
            cb = codes.download_at(c * I * T, np.uint8, I * T).reshape(I, T)
            x = lut[cb & 3] + 1j * lut[(cb >> 2) & 3]
            ref[:, k] = (x @ x.conj().T)[iu]
        err = {k: nrel(v, ref) for k, v in vis_sample.items()}
        # the exact engines against the FP64 brute force rounded to FP32: units in the last place
        r32 = ref.astype(np.complex64).view(np.int32).astype(np.int64)
        ulp = {k: np.abs(vis_sample[k].view(np.int32).astype(np.int64) - r32) for k in ("fp4_2bit", "fp8_2bit")}
        err["exact_vs_fp64_rounded"] = {k: {"max_ulp": int(u.max()), "equal_frac": float(np.mean(u == 0))}
                                        for k, u in ulp.items()}
        err["sample"] = f"channels {sample_ch}, all {P} products, all {T} spectra, vs FP64"
        log(f"[bench] large-N errors: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()
                                                    if k not in ("sample", "exact_vs_fp64_rounded"))
            + f"; exact engines vs FP64 rounded to FP32: {err['exact_vs_fp64_rounded']}")
    if not args.no_cpu:
        # FP32 restatement of the X-engine on a sample of channels, all inputs and spectra
        O = args.oracle
        nth = cpu_threads()[0]
        cs, tsmp = 128, min(T, 512)
        X = np.empty((tsmp, I, cs), np.complex64)          # [t][i][c]: the generic core's loop order
        for k in range(cs):
            cb = codes.download_at(k * I * T, np.uint8, I * T).reshape(I, T)[:, :tsmp].T
            X[:, :, k] = lut[cb & 3] + 1j * lut[(cb >> 2) & 3]
        dt = cpu_median(lambda: O.xcorr_tic_f32(X, nth))
        cm = float(P) * cs * tsmp
        out["cpu_baseline"] = cpu_obj(cm / dt, "complex MAC/s",
                                      f"{cs} of {C} channels x {tsmp} of {T} spectra x {I} inputs (FP32 X-engine), "
                                      f"{dt:.2f} s per run", {"input_gsps": I * cs * tsmp / dt / 1e9})
        log(f"[bench] large-N CPU baseline: {cm / dt:.3e} CMAC/s on {nth} threads")
        del X
    codes.free()
    vis.free()
    flop = 8.0 * cmac
    out["ms"] = res
    out["kernels"] = knames
    out["mfma_busy"] = {k: measured_busy(n, "mfma") for k, n in knames.items()}
    out["complex_mac_per_s"] = {k: cmac / (v * 1e-3) for k, v in res.items()}
    out["tflops_equiv"] = {k: flop / (v * 1e-3) / 1e12 for k, v in res.items()}
    out["speedup_vs_fp32"] = ({k: res["f32"] / v for k, v in res.items() if k != "f32"}
                              if "f32" in res else None)
    out["err_vs_fp64"] = err
    out["note"] = ("fp8_2bit / fp4_2bit run three e4m3 / e2m1 GEMMs (S, S+T, T of X = S + (H-1)T), so "
                   "their MFMA work is 3x the algorithmic count; tflops_equiv counts 8*P*C*T")
    return out


def spectrometer(fxlib, ctx, args):
    """The software spectrometer's workload: 16 interleaved 2-bit subband
    channels, 2^28 samples in all (2^24 per channel, 64 MiB packed, resident
    in HBM), 32768-point real FFT per channel, power accumulated over the
    whole stream (512 FFTs per channel, one spectrum of 16384+1 bins each),
    window off, no overlap (fx_spectrometer_ex)."""
    nch, n = 16, 32768
    ntime = (1 << 28) // nch
    K = ntime // n
    dw = ctx.alloc(ntime * nch // 4)
    rng = np.random.default_rng(args.seed + 9)
    amp = np.zeros(nch)
    freq = np.zeros(nch)
    for c in rng.choice(nch, 3, replace=False):
        amp[c], freq[c] = 0.2, rng.uniform(0.05, 0.45)
    ctx.synth_multichannel(args.seed + 9, nch, amp, freq, 0.9816, ntime, dw)
    dout = ctx.alloc(nch * (n // 2 + 1) * 4)
    lib = fxlib.load()
    ns = fxlib.C.c_long()

    def run():
        fxlib.check(lib.fx_spectrometer_ex(ctx.handle, dw.ptr, ntime, nch, n, 1, K, fxlib.FX_WIN_NONE, 0, 1,
                                           dout.ptr, fxlib.C.byref(ns)))

    run()
    ctx.sync()
    e = [ctx.event() for _ in range(2)]
    reps = max(3, args.steps)
    ctx.record(e[0])
    for _ in range(reps):
        run()
    ctx.record(e[1])
    ctx.sync()
    ms = ctx.elapsed_ms(e[0], e[1]) / reps
    nfft = nch * K
    out = {"workload": "16 channels x 2^24 2-bit samples, 32768-pt real FFT, power over the whole stream",
           "ms": ms, "ffts_per_s": nfft / (ms * 1e-3), "gsamples_per_s": (1 << 28) / (ms * 1e-3) / 1e9,
           "fft_tflops": 2.5 * n * np.log2(n) * nfft / (ms * 1e-3) / 1e12, "spectra": int(ns.value)}
    log(f"[bench] spectrometer: {ms:.2f} ms for 2^28 samples ({out['gsamples_per_s']:.0f} GS/s, "
        f"{out['ffts_per_s']:.3g} FFT/s)")
    if not args.no_cpu:
        O = args.oracle
        nth = cpu_threads()[0]
        words = dw.download(np.uint32)
        got = dout.download(np.float32, shape=(1, nch, n // 2 + 1))
        ref = O.spectrometer_ex_f32(words, ntime, nch, n, 1, K, O.WIN_NONE, 0, 1, nthreads=nth)
        out["err_vs_cpu_fp32"] = max(nrel(got[0, c], ref[0, c]) for c in range(nch))
        dt = cpu_median(lambda: O.spectrometer_ex_f32(words, ntime, nch, n, 1, K, O.WIN_NONE, 0, 1, nthreads=nth))
        out["cpu_baseline"] = cpu_obj((1 << 28) / dt / 1e9, "input GS/s", "the whole workload",
                                      {"ms": dt * 1e3, "ffts_per_s": nfft / dt})
        out["gpu_over_cpu"] = dt * 1e3 / ms
        log(f"[bench] spectrometer CPU baseline: {dt * 1e3:.0f} ms on {nth} threads")
    dw.free()
    dout.free()
    return out


def unpack(fxlib, ctx, args):
    """2-bit unpack to FP32 through the level table (fx_unpack2b): 2^30
    samples of one channel (256 MiB packed -> 4 GiB of floats), and 2^26 time
    steps of 16 interleaved channels (same bytes, channel-major out).
    Algorithmic bytes: 0.25 B read + 4 B written per sample."""
    ns = 1 << 30
    dw = ctx.alloc(ns // 4)
    fxlib.check(fxlib.load().fx_synth_station(ctx.handle, args.seed + 8, 0, 0.0, 0, 0.0, 0.0, 0.9816, ns,
                                              dw.ptr))
    dout = ctx.alloc(ns * 4)
    out = {"workload": "2^30 2-bit samples -> FP32: 1 channel, and 16 interleaved channels x 2^26 time steps",
           "algorithmic": "0.25 B in + 4 B out per sample"}
    e = [ctx.event() for _ in range(2)]
    reps = max(3, args.steps)
    for name, nch in (("1ch", 1), ("16ch", 16)):
        ntime = ns // nch
        ctx.unpack2b(dw, ntime, nch, dout)
        ctx.sync()
        ctx.record(e[0])
        for _ in range(reps):
            ctx.unpack2b(dw, ntime, nch, dout)
        ctx.record(e[1])
        ctx.sync()
        ms = ctx.elapsed_ms(e[0], e[1]) / reps
        gbps = ns * 4.25 / (ms * 1e-3) / 1e9
        out[name] = {"ms": ms, "gsamples_per_s": ns / (ms * 1e-3) / 1e9, "gbps": gbps, "frac": gbps / PEAK["hbm"]}
    log(f"[bench] unpack: 1ch {out['1ch']['ms']:.2f} ms ({out['1ch']['gbps']:.0f} GB/s), "
        f"16ch {out['16ch']['ms']:.2f} ms ({out['16ch']['gbps']:.0f} GB/s)")
    dw.free()
    dout.free()
    return out


def products(fxlib, ctx, dvis, P, C, T, args):
    """Lag spectra and amplitude/phase of the dumped visibilities of the
    mid-size correlator (zerocorr.c:620-645 downstream products): one batched
    backward FFT of P rows x C channels, then |.| and arg of the lags."""
    n = P * C
    dl = ctx.alloc(n * 8)
    da, dp = ctx.alloc(n * 4), ctx.alloc(n * 4)
    scale = 1.0 / T
    ctx.lags(dvis, P, C, scale, dl)
    ctx.amp_phase(dl, n, 1.0, da, dp)
    ctx.sync()
    reps = max(3, args.steps)
    e = [ctx.event() for _ in range(3)]
    ctx.record(e[0])
    for _ in range(reps):
        ctx.lags(dvis, P, C, scale, dl)
    ctx.record(e[1])
    for _ in range(reps):
        ctx.amp_phase(dl, n, 1.0, da, dp)
    ctx.record(e[2])
    ctx.sync()
    lag_ms = ctx.elapsed_ms(e[0], e[1]) / reps
    ap_ms = ctx.elapsed_ms(e[1], e[2]) / reps
    out = {"workload": f"{P} products x {C} channels: lag spectra (backward FFT, 1/n_spectra) then "
                       f"amplitude and phase",
           "lags_ms": lag_ms, "amp_phase_ms": ap_ms,
           "lags_gbps": 2.0 * n * 8 / (lag_ms * 1e-3) / 1e9,
           "lags_fft_tflops": 5.0 * C * np.log2(C) * P / (lag_ms * 1e-3) / 1e12,
           "amp_phase_gbps": n * (8 + 8) / (ap_ms * 1e-3) / 1e9}
    log(f"[bench] products: lags {lag_ms:.2f} ms ({out['lags_gbps']:.0f} GB/s), "
        f"amp/phase {ap_ms:.2f} ms ({out['amp_phase_gbps']:.0f} GB/s)")
    for b in (dl, da, dp):
        b.free()
    return out


def front_end(fxlib, ctx, args):
    """VDIF front end: 8 threads x 4000 frames/s x 4 s of 8032-byte frames
    (8000-byte 2-bit payloads, 4 channels; about 1 GB) resident in HBM, placed
    into per-thread rows by fx_frames_place; then the same bytes streamed from
    a file through the double-buffered ingest (page-cache reads, pinned
    buffers, PCIe copies) and placed chunk by chunk."""
    import tempfile
    nth, fps, nsec, pb, hb = 8, 4000, 4, 8000, 32
    fb = pb + hb
    nf = nth * fps * nsec
    rng = np.random.default_rng(args.seed + 5)
    frames = np.empty((nf, fb), np.uint8)
    frames[:, hb:] = rng.integers(0, 256, (nf, pb), dtype=np.uint8)
    idx = np.arange(nf, dtype=np.int64)
    sec0 = 50_000_000
    w = np.zeros((nf, hb // 4), "<u4")
    w[:, 0] = sec0 + idx // (fps * nth)
    w[:, 1] = ((idx // nth) % fps) | (40 << 24)
    w[:, 2] = (fb // 8) | (2 << 24)
    w[:, 3] = ((idx % nth) << 16) | (1 << 26)
    frames[:, :hb] = w.view(np.uint8).reshape(nf, hb)
    d_raw = ctx.to_device(frames.ravel())
    nslots, pw = fps * nsec, pb // 4
    stride = nslots * pw
    d_map = ctx.to_device(np.arange(nth, dtype=np.int32))
    d_out = ctx.alloc(4 * nth * stride)
    d_valid, d_counts = ctx.alloc(8 * nth * nslots), ctx.alloc(32)

    def place(ptr, n, clear):
        ctx.frames_place(fxlib.FX_FMT_VDIF, ptr, n, fb, hb, d_map, nth, nth, sec0, 0, fps, nslots, d_out, stride,
                         d_valid, d_counts, clear=clear)

    place(d_raw, nf, True)
    ctx.sync()
    counts = d_counts.download(np.uint64)
    probe = [(3, 777), (7, nslots - 1), (0, 0)]
    ok = int(counts[0]) == nf and int(counts[1]) == 0 and all(
        np.array_equal(d_out.download_at(4 * (r * stride + s * pw), np.uint32, pw),
                       frames[s * nth + r, hb:].view("<u4")) for r, s in probe)
    reps = max(3, args.steps)
    e = [ctx.event() for _ in range(2)]
    ctx.record(e[0])
    for _ in range(reps):
        place(d_raw, nf, True)
    ctx.record(e[1])
    ctx.sync()
    ms = ctx.elapsed_ms(e[0], e[1]) / reps
    alg = nf * (fb + pb + 8.0)
    out = {"workload": f"VDIF {nth} threads x {fps} frames/s x {nsec} s, {fb}-byte frames ({nf} frames, "
                       f"{nf * fb / 1e9:.2f} GB) resident in HBM -> per-thread rows",
           "place_ms": ms, "frames_per_s": nf / (ms * 1e-3),
           "payload_gbps": nf * pb / (ms * 1e-3) / 1e9,
           "roofline": {"bound": "hbm", "achieved": alg / (ms * 1e-3) / 1e9, "peak": PEAK["hbm"], "unit": "GB/s",
                        "frac": alg / (ms * 1e-3) / 1e9 / PEAK["hbm"],
                        "algorithmic": "per frame: framebytes read + payload written + 8 B range"},
           "parity_probe_ok": bool(ok)}
    tr = measured_traffic("k_frames_place<uint4, 8>")
    out["roofline"]["traffic"] = None if tr is None else tr["bytes"]
    if tr is not None:
        out["roofline"]["traffic_source"] = tr["source"]
    d_raw.free()
    # the same bytes from a file through the double-buffered ingest (half the
    # stream, so the temporary file stays small)
    nfi = nf // 2
    chunk_frames = 8192
    with tempfile.NamedTemporaryFile(prefix="fx_vdif_", suffix=".vdif", dir=os.environ.get("TMPDIR", "/tmp")) as tf:
        frames[:nfi].tofile(tf.name)
        with open(tf.name, "rb") as fh:          # warm the page cache
            while fh.read(64 << 20):
                pass
        ctx.sync()
        t0 = time.time()
        g = fxlib.Ingest(ctx, tf.name, chunk_frames * fb, nbuf=3)
        first = True
        tot = 0
        for p, n in g:
            place(p, n // fb, first)
            first = False
            tot += n
        ctx.sync()
        dt = time.time() - t0
        g.close()
    counts = d_counts.download(np.uint64)
    out["ingest"] = {"bytes": tot, "seconds": dt, "gbps": tot / dt / 1e9, "chunk_bytes": chunk_frames * fb,
                     "nbuf": 3, "placed": int(counts[0]),
                     "note": "file in the page cache -> pinned host buffers -> PCIe -> placement; "
                             "includes the host read"}
    log(f"[bench] front end: place {ms:.2f} ms for {nf} frames ({out['roofline']['achieved']:.0f} GB/s, "
        f"frac {out['roofline']['frac']:.2f}), ingest {out['ingest']['gbps']:.1f} GB/s")
    for b in (d_map, d_out, d_valid, d_counts):
        b.free()
    return out


def spectrometer_app(fxlib, args):
    """Spectrometer application: two 256 MiB raw files of single-channel
    2-bit samples (2^30 samples each, 16 MHz sampling: 67 s of data) through
    fx_swspectrometer_run -- 1024-point FFTs with 50% overlap (the reference's
    spectrometer benchmark shape), Hann window, 1 s integrations (31249 FFTs
    per spectrum), auto spectra of both inputs and their cross spectrum to the
    binary sinks.  FFT/s counts the overlapped FFTs of one input."""
    import tempfile
    import tasklib as TL
    n, ov, fs = 1024, 2, 16e6
    K = int(fs / (n // ov)) - 1                 # (K-1)*hop + n samples = 1 s
    ntime = 1 << 30
    rng = np.random.default_rng(args.seed + 7)
    tdir = os.environ.get("TMPDIR", "/tmp")
    out = {"workload": f"2 raw files x {ntime} single-channel 2-bit samples ({ntime / fs:.0f} s at 16 MHz), "
                       f"{n}-pt FFT, 50% overlap, Hann, {K} FFTs per 1 s spectrum, autos + cross spectrum"}
    with tempfile.TemporaryDirectory(dir=tdir) as d:
        block = rng.integers(0, 2**32, 1 << 22, dtype=np.uint64).astype(np.uint32)   # 16 MiB
        paths = []
        for i in range(2):
            p = os.path.join(d, f"in{i}.raw")
            with open(p, "wb") as fh:
                for k in range(ntime // 16 // block.size):
                    np.roll(block, 4099 * (k + 17 * i)).tofile(fh)
            paths.append(p)
        for p in paths:                            # page cache
            with open(p, "rb") as fh:
                while fh.read(64 << 20):
                    pass
        s = TL.SwspecSettings()
        c = s.core
        c.fft_points, c.fft_overlap_factor, c.core_overlapped_ffts = n, ov, K
        c.window, c.bits_per_sample, c.source_channels, c.samplingfreq = fxlib.FX_WIN_HANN, 2, 1, fs
        s.format = fxlib.FX_FMT_RAW
        s.use_channel_file[0] = s.use_channel_file[1] = 1
        s.calc_xpol, s.spectra_per_chunk = 1, 8
        s.basefilename[0] = os.path.join(d, "out1").encode()
        s.basefilename[1] = os.path.join(d, "out2").encode()
        st = TL.SwspecStats()
        rc = TL.load().fx_swspectrometer_run(s, paths[0].encode(), paths[1].encode(), st)
        if rc != 0:
            raise SystemExit(f"bench: spectrometer application failed: {fxlib.load().fx_last_error()}")
    nfft = st.spectra * K
    out.update({"spectra": st.spectra, "seconds": st.seconds, "ffts_per_s_per_input": nfft / st.seconds,
                "realtime_factor": st.realtime_factor,
                "note": "files in the page cache; host reads, PCIe copies, device compute and sink writes "
                        "all inside the timed loop"})
    log(f"[bench] spectrometer app: {st.spectra} spectra in {st.seconds:.2f} s "
        f"({nfft / st.seconds:.3g} FFT/s per input, {st.realtime_factor:.1f}x realtime)")
    return out


def task_chains(args):
    """Task runtime: a five-task next_task chain (1024-point FFT -> complex
    multiply -> complex MAC, a 2-channel unpack of 4096 samples, a DiFX
    process task fed by the FFT output) executed repeatedly through
    celldsp_executeTask, per task (stream synchronization after each) and as
    one replayed hipGraph (CDSP_BEHAV_GRAPH).  Host wall time per chain."""
    import tasklib as TL
    lib = TL.load()
    lib.celldsp_start()
    n, nch = 1024, 256
    rng = np.random.default_rng(args.seed + 6)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    X, Y, acc = (np.zeros(n, np.complex64) for _ in range(3))
    m = np.exp(1j * rng.uniform(0, 2 * np.pi, n)).astype(np.complex64)
    words = rng.integers(0, 2**32, 512, dtype=np.uint64).astype(np.uint32)
    f = np.zeros(8192, np.float32)
    mult = np.exp(1j * rng.uniform(0, 2 * np.pi, nch + 1)).astype(np.complex64)
    fo, cfo, acc2 = (np.zeros(nch + 1, np.complex64) for _ in range(3))
    t = [TL.CellDSPTaskClass(c) for c in (TL.CELLDSP_TASK_FFT_EXEC_DYN, TL.CELLDSP_TASK_MULC,
                                          TL.CELLDSP_TASK_MAC, TL.CELLDSP_TASK_UNPACK,
                                          TL.CELLDSP_TASK_DIFXPROCESS)]
    t[0].setFFTparams(n)
    t[0].addInVec(x); t[0].addOutVec(X)
    t[1].addInVec(X); t[1].addInVec(m); t[1].addOutVec(Y)
    t[2].addInVec(Y); t[2].addInVec(m); t[2].addOutVec(acc)
    t[3].task.subcontext.Unpack.nchan = 2
    t[3].addInVec(words); t[3].addOutVec(f)
    d = t[4].task.subcontext.Difx
    d.nfft, d.numchannels, d.useUpperOrLower = n, nch, 1
    t[4].addInVec(X); t[4].addInVec(mult)
    t[4].addOutVec(fo); t[4].addOutVec(cfo); t[4].addOutVec(acc2)
    for a, b in zip(t, t[1:]):
        a.setNext(b)
    out = {"workload": "5-task chain: FFT 1024 -> MULC -> MAC, UNPACK 2x4096, DIFXPROCESS 1024/256",
           "unit": "us per chain (host wall time)"}
    reps = 200
    for name, beh in (("per_task", TL.CDSP_BEHAV_DEFAULT), ("graph", TL.CDSP_BEHAV_GRAPH)):
        lib.celldsp_setBehaviour(beh)
        for _ in range(5):
            lib.celldsp_executeTask(t[0].getTask())
        t0 = time.perf_counter()
        for _ in range(reps):
            if lib.celldsp_executeTask(t[0].getTask()) != 0:
                raise SystemExit("bench: task chain failed")
        out[name] = (time.perf_counter() - t0) / reps * 1e6
    out["graph_replays"] = int(lib.celldsp_graphReplays())
    lib.celldsp_setBehaviour(TL.CDSP_BEHAV_DEFAULT)
    lib.celldsp_stop()
    log(f"[bench] task chain: per task {out['per_task']:.0f} us, graph {out['graph']:.0f} us")
    return out


def batch_draw(seed, nsub, nsamp=1 << 22):
    """The uneven-subband draw: per subband an FFT length from
    {256, 1024, 4096, 16384} and 8..48 stations (x2 polarizations); returns
    [(nfft, ninp, nspectra)] for nsamp samples per input."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nsub):
        n = int(rng.choice([256, 1024, 4096, 16384]))
        S = int(rng.integers(8, 49))
        out.append((n, 2 * S, nsamp // n))
    return out


def batched(fxlib, ctx, args):
    """32 subbands with FFT lengths drawn from {256,1024,4096,16384}, station
    counts from 8..48 (x2 pol), rho = 0.03, 2^22 samples per input, FP64
    phase-cal of the 10 kHz-offset 1 MHz comb on every input (one comb tooth
    injected per subband): one hipGraph launch of the autotuned FP16 batch vs
    the same subbands one launch each in plain FP32."""
    import tempfile
    nsub = args.batch_subbands
    nsamp = 1 << 22
    subs = []
    bufs = []
    tot_cmac = 0.0
    tot_samp = 0.0
    for k, (n, I, T) in enumerate(batch_draw(args.seed + 4, nsub, nsamp)):
        nw = nsamp // 16
        w = ctx.alloc(I * nw * 4)
        for i in range(I):
            tone = 0.2 if i == 0 else 0.0
            fxlib.check(fxlib.load().fx_synth_station(ctx.handle, args.seed + 100 + k, i, 0.03, i % 7, tone,
                                                      (10e3 + 3e6) / 16e6, 0.9816, nsamp,
                                                      w.ptr.value + i * nw * 4))
        bufs.append(w)
        subs.append({"nfft": n, "ninp": I, "nspectra": T, "words": w, "stride_words": nw,
                     "models": None, "pcal": (8e6, 1e6, 10e3)})
        tot_cmac += (I * (I + 1) / 2) * (n / 2) * T
        tot_samp += I * nsamp
    ctx.sync()
    tune = os.path.join(tempfile.gettempdir(), f"fx_tune_{os.getpid()}.txt")
    t0 = time.time()
    p16 = fxlib.Batch(ctx, subs, fxlib.FX_PREC_F16, use_graph=fxlib.FX_BATCH_PERSISTENT, tune_file=tune)
    log(f"[bench] batch: {nsub} subbands, persistent table autotuned + captured in {time.time() - t0:.1f}s")
    ph16 = fxlib.Batch(ctx, subs, fxlib.FX_PREC_F16, use_graph=fxlib.FX_BATCH_PHASED, tune_file=tune)
    t0 = time.time()
    g16 = fxlib.Batch(ctx, subs, fxlib.FX_PREC_F16, use_graph=fxlib.FX_BATCH_GRAPH, tune_file=tune)
    log(f"[bench] batch: per-subband graph autotuned + captured in {time.time() - t0:.1f}s")
    e32 = fxlib.Batch(ctx, subs, fxlib.FX_PREC_F32, use_graph=False)

    def timed(b, reps):
        b.run()
        b.sync()
        e0, e1 = ctx.event(), ctx.event()
        ctx.record(e0)
        for _ in range(reps):
            b.run()
        ctx.record(e1)
        ctx.sync()
        return ctx.elapsed_ms(e0, e1) / reps

    msp = timed(p16, args.steps)
    msph = timed(ph16, args.steps)
    ms16 = timed(g16, args.steps)
    ms32 = timed(e32, max(1, args.steps - 1))
    p16.sync()
    nitems, nblocks, lds = p16.persist_info()
    log(f"[bench] batch: fp16 persistent {msp:.2f} ms ({nitems} items on {nblocks} workgroups), phased "
        f"{msph:.2f} ms, per-subband graph {ms16:.2f} ms, fp32 one launch each {ms32:.2f} ms")
    err = None
    nth = cpu_threads()[0]
    frac = 64
    if not args.no_check:
        # every subband against the oracle on its first 1/64 of the integration
        # (at least 8 spectra): the same F -> X chain, a shorter sum
        O = args.oracle
        err = {"sample": f"every subband, its first 1/{frac} of the spectra (>= 8)"}
        sub16 = fxlib.Batch(ctx, [dict(s, nspectra=max(8, s["nspectra"] // frac)) for s in subs],
                            fxlib.FX_PREC_F16, use_graph=fxlib.FX_BATCH_PERSISTENT, tune_file=tune)
        sub32 = fxlib.Batch(ctx, [dict(s, nspectra=max(8, s["nspectra"] // frac)) for s in subs],
                            fxlib.FX_PREC_F32, use_graph=False)
        for b in (sub16, sub32):
            b.run()
            b.sync()
        worst16 = worst32 = 0.0
        for k, s in enumerate(subs):
            ts = max(8, s["nspectra"] // frac)
            st = [s["words"].download_at(i * s["stride_words"] * 4, np.uint32, ts * s["nfft"] // 16)
                  for i in range(s["ninp"])]
            r64 = O.corr_f64(st, s["nfft"], ts, None, nthreads=nth)
            r32 = O.corr_f32(st, s["nfft"], ts, None, nthreads=nth)
            worst16 = max(worst16, nrel(sub16.visibilities(k), r64))
            worst32 = max(worst32, nrel(sub32.visibilities(k), r32))
        err["f16_vs_fp64_worst"] = worst16
        err["f32_vs_fp32_oracle_worst"] = worst32
        sub16.close()
        sub32.close()
        log(f"[bench] batch errors (all {nsub} subbands): f16 vs FP64 worst {worst16:.2e}, "
            f"f32 vs FP32 oracle worst {worst32:.2e}")
    cpu = None
    if not args.no_cpu:
        O = args.oracle
        streams = []
        for s in subs:
            ts = max(8, s["nspectra"] // frac)
            streams.append([s["words"].download_at(i * s["stride_words"] * 4, np.uint32, ts * s["nfft"] // 16)
                            for i in range(s["ninp"])])

        def cpu_all():
            for s, st in zip(subs, streams):
                O.corr_f32(st, s["nfft"], max(8, s["nspectra"] // frac), None, nthreads=nth)

        dt = cpu_median(cpu_all)
        cm = sum((s["ninp"] * (s["ninp"] + 1) / 2) * (s["nfft"] / 2) * max(8, s["nspectra"] // frac) for s in subs)
        smp = sum(s["ninp"] * s["nfft"] * max(8, s["nspectra"] // frac) for s in subs)
        cpu = cpu_obj(cm / dt, "complex MAC/s", f"all {nsub} subbands, 1/{frac} of each integration (>= 8 spectra), "
                      f"FP32 F -> X, no phase-cal, {dt:.2f} s per run", {"input_gsps": smp / dt / 1e9})
        log(f"[bench] batch CPU baseline: {cm / dt:.3e} CMAC/s on {nth} threads")
    tuning = [g16.tuning(k) for k in range(nsub)]
    ptuning = [p16.persist_tuning(k) for k in range(nsub)]
    for b in (p16, ph16, g16, e32):
        b.close()
    for w in bufs:
        w.free()
    try:
        os.remove(tune)
    except OSError:
        pass
    return {"workload": f"{nsub} subbands, N in {{256,1024,4096,16384}}, 16-96 inputs, rho=0.03, "
                        f"2^22 samples per input, FP64 phase-cal of every input",
            "subbands": [{"nfft": s["nfft"], "inputs": s["ninp"], "spectra": s["nspectra"],
                          "graph_sub_spectra": t[0], "graph_xtile": t[1], "persistent_natural_rows": q[0],
                          "persistent_channels_per_x_item": q[1], "persistent_f_groups_per_item": q[2],
                          "persistent_time_slices": q[3], "persistent_xtile": q[4]}
                         for s, t, q in zip(subs, tuning, ptuning)],
            "persistent_launch": {"items": nitems, "workgroups": nblocks, "lds_bytes": lds},
            "ms": {"f16_persistent": msp, "f16_phased": msph, "f16_graph": ms16, "f32_one_launch_each": ms32},
            "complex_mac_per_s": {"f16_persistent": tot_cmac / (msp * 1e-3), "f16_graph": tot_cmac / (ms16 * 1e-3),
                                  "f32_one_launch_each": tot_cmac / (ms32 * 1e-3)},
            "input_gsps": {"f16_persistent": tot_samp / (msp * 1e-3) / 1e9, "f16_graph": tot_samp / (ms16 * 1e-3) / 1e9,
                           "f32_one_launch_each": tot_samp / (ms32 * 1e-3) / 1e9},
            "speedup": ms32 / msp, "complex_macs": tot_cmac, "err": err, "cpu_baseline": cpu}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--stations", type=int, default=64)
    ap.add_argument("--nfft", type=int, default=8192)
    ap.add_argument("--spectra", type=int, default=65536)
    ap.add_argument("--chunk", type=int, default=16384,
                    help="spectra per chunk of the mid-size correlator (16384: 34 GB of FP16 channels, 69 GB FP32)")
    ap.add_argument("--rho", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=20240229)
    ap.add_argument("--large-inputs", type=int, default=512)
    ap.add_argument("--large-channels", type=int, default=1024)
    ap.add_argument("--large-spectra", type=int, default=8192)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-fp8", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-front-end", action="store_true")
    ap.add_argument("--no-tasks", action="store_true")
    ap.add_argument("--batch-subbands", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU budget of the headline workload's CPU baseline")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="instead of the benchmark, run the plain copy kernel over MIB MiB "
                         "(HBM bandwidth check and traffic-counter calibration)")
    ap.add_argument("--large-only", action="store_true",
                    help="instead of the benchmark, run only the large-N X-engine section")
    ap.add_argument("--batch-only", action="store_true",
                    help="instead of the benchmark, run only the batch of uneven subbands")
    ap.add_argument("--front-end-only", action="store_true",
                    help="instead of the benchmark, run only the front-end, task-chain and spectrometer-application sections")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the F-X chain shards by time chunk with an RCCL reduce of the visibilities, "
              "which is designed but not built)", file=sys.stderr)
        sys.exit(2)

    import fxlib

    # the CPU restatement, built -O3 -march=native for this host: the CPU
    # baseline and the error checks' oracle
    args.oracle = None
    if not (args.no_cpu and args.no_check):
        import tempfile
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        ndir = tempfile.mkdtemp(prefix="fx_oracle_")
        O.build_native(os.path.join(ndir, "liboracle_fx_native.so"))
        args.oracle = O

    if args.copy_probe:
        copy_probe(fxlib, args)
        return
    if args.large_only:
        ctx = fxlib.Context(0)
        print(json.dumps({"large_n_xengine": large_n(fxlib, ctx, args)}), flush=True)
        ctx.close()
        return
    if args.batch_only:
        ctx = fxlib.Context(0)
        print(json.dumps({"batched_subbands": batched(fxlib, ctx, args)}), flush=True)
        ctx.close()
        return
    if args.front_end_only:
        ctx = fxlib.Context(0)
        print(json.dumps({"unpack": unpack(fxlib, ctx, args), "spectrometer": spectrometer(fxlib, ctx, args),
                          "front_end": front_end(fxlib, ctx, args),
                          "task_chains": task_chains(args),
                          "spectrometer_app": spectrometer_app(fxlib, args)}), flush=True)
        ctx.close()
        return

    I, N, C, T, tau, models = workload(args)
    P = I * (I + 1) // 2
    nwords = T * N // 16            # packed words per input
    chunk = min(args.chunk, T)
    nchunks = (T + chunk - 1) // chunk
    log(f"[bench] workload: I={I} N={N} C={C} T={T} P={P} chunk={chunk}")

    ctx = fxlib.Context(0)
    # ---- synthetic input resident in HBM --------------------------------
    t0 = time.time()
    dw = ctx.alloc(I * nwords * 4)
    base = dw.ptr.value
    lib = fxlib.load()
    for i in range(I):
        s = i // 2
        fxlib.check(lib.fx_synth_station(ctx.handle, args.seed, i, args.rho, int(tau[s]), 0.0, 0.0,
                                         0.9816, T * N, base + i * nwords * 4))
    ctx.sync()
    log(f"[bench] generated {I * nwords * 4 / 2**30:.1f} GiB of 2-bit input in {time.time() - t0:.1f}s")
    dvis = ctx.alloc(P * C * 8)

    def run_path(prec, esz, steps, warmup):
        """Whole steps with events around the F stage and the X-engine of chunk
        0; returns ms/step and their durations (the F stage of a chunk is one
        launch of the direct engine-layout transform kernel where it applies)."""
        corr = fxlib.Correlator(ctx, I, N, chunk, prec, use_phase=True)
        corr.set_phase_model(models)
        dchan = ctx.alloc(C * chunk * I * esz)
        nf = corr.fstage_launches(chunk, prec)
        fname = corr.fstage_name(chunk, prec)
        xname = corr.xengine_name(chunk, dchan)
        evf = [ctx.event(), ctx.event()]
        evx = [ctx.event(), ctx.event()]

        def step():
            corr.reset()
            for k in range(nchunks):
                nt = min(chunk, T - k * chunk)
                if k == 0:
                    ctx.record(evf[0])
                corr.fstage(dw, nwords, k * chunk, nt, dchan, prec)
                if k == 0:
                    ctx.record(evf[1])
                    ctx.record(evx[0])
                corr.xengine(dchan, nt)
                if k == 0:
                    ctx.record(evx[1])
            corr.dump(dvis)

        for _ in range(warmup):
            step()
        ctx.sync()
        e0, e1 = ctx.event(), ctx.event()
        ctx.record(e0)
        for _ in range(steps):
            step()
        ctx.record(e1)
        ctx.sync()
        ms = ctx.elapsed_ms(e0, e1) / steps
        # per-launch durations in chunk 0 of the last timed step (events are
        # re-recorded each step)
        f = ctx.elapsed_ms(evf[0], evf[1])
        x = ctx.elapsed_ms(evx[0], evx[1])
        dchan.free()
        corr.close()
        return ms, {"fstage": f, "xengine": x, "fstage_launches": nf, "fstage_kernel": fname,
                    "xengine_kernel": xname}

    paths = {}
    ms16, k16 = run_path(fxlib.FX_PREC_F16, 4, args.steps, args.warmup)
    paths["f16"] = (ms16, k16)
    log(f"[bench] fp16 path: {ms16:.2f} ms/step; per chunk " +
        ", ".join(f"{k} {v}" if isinstance(v, (int, str)) else f"{k} {v:.2f} ms" for k, v in k16.items()))
    if not args.no_fp32:
        paths["f32"] = run_path(fxlib.FX_PREC_F32, 8, max(1, args.steps - 1), 1)
        log(f"[bench] fp32 path: {paths['f32'][0]:.2f} ms/step")
    if not args.no_fp8:
        paths["fp8"] = run_path(fxlib.FX_PREC_FP8, 2, max(1, args.steps - 1), 1)
        log(f"[bench] fp8 path: {paths['fp8'][0]:.2f} ms/step")

    cmac = float(P) * C * T
    value = cmac / (ms16 * 1e-3)
    input_gsps = I * float(N) * T / (ms16 * 1e-3) / 1e9

    # ---- error vs the FP64 oracle on a bounded sub-run ------------------------
    err = None
    O = args.oracle
    if not args.no_check:
        # a multiple of every engine's stage depth, so the error is measured on
        # the same X-engine kernel the timed steps run
        te = 32
        streams = [dw.download_at(i * nwords * 4, np.uint32, te * N // 16) for i in range(I)]
        got = {}
        err_kernels = {}
        for name, prec in (("f16", fxlib.FX_PREC_F16), ("f32", fxlib.FX_PREC_F32), ("fp8", fxlib.FX_PREC_FP8)):
            c = fxlib.Correlator(ctx, I, N, te, prec, use_phase=True)
            c.set_phase_model(models)
            err_kernels[name] = c.xengine_name(te)
            c.process(dw, nwords, 0, te)
            got[name] = c.visibilities()
            c.close()
        nth = cpu_threads()[0]
        t0 = time.time()
        ref64 = O.corr_f64(streams, N, te, models, nthreads=nth)
        log(f"[bench] FP64 oracle on {te} spectra in {time.time() - t0:.1f}s")
        err = {k: nrel(v, ref64) for k, v in got.items()}
        err["sample"] = f"first {te} spectra of all {I} inputs, all {P} products x {C} channels"
        err["xengine_kernels"] = err_kernels
        for name, (_ms, kt) in paths.items():
            if kt["xengine_kernel"] != err_kernels[name]:
                raise SystemExit(f"bench: the {name} error sample ran {err_kernels[name]}, "
                                 f"the timed steps {kt['xengine_kernel']}")
        log(f"[bench] normwise error vs FP64 oracle: " +
            ", ".join(f"{k} {v:.2e}" for k, v in err.items() if isinstance(v, float)))
        del ref64, got

    # ---- CPU baseline (oracle FP32 generic-C chain, the host's CPU share) ----
    cpu = None
    if not args.no_cpu:
        nth = cpu_threads()[0]
        # a bounded sample: a probe run on 64 spectra sizes the sample so one
        # run takes about a quarter of the budget (four runs in all)
        ts = min(T, 64)
        streams = [dw.download_at(i * nwords * 4, np.uint32, ts * N // 16) for i in range(I)]
        t0 = time.time()
        O.corr_f32(streams, N, ts, models, nthreads=nth)
        dt = time.time() - t0
        while ts * 2 <= T and dt * 2 < args.cpu_seconds / 4:
            ts *= 2
            dt *= 2
        streams = [dw.download_at(i * nwords * 4, np.uint32, ts * N // 16) for i in range(I)]
        dt = cpu_median(lambda: O.corr_f32(streams, N, ts, models, nthreads=nth))
        cpu = cpu_obj(P * C * ts / dt, "complex MAC/s",
                      f"{ts} of {T} spectra x {I} inputs x {C} channels, FP32 F -> X with the FP64 phase model, "
                      f"{dt:.2f} s per run", {"input_gsps": I * N * ts / dt / 1e9})
        log(f"[bench] CPU baseline: {cpu['value']:.3e} CMAC/s on {nth} threads ({ts} spectra, {dt:.2f}s)")
    dw.free()
    prod = products(fxlib, ctx, dvis, P, C, T, args)
    dvis.free()
    mini = mini_correlator(fxlib, ctx, args, O)
    unp = unpack(fxlib, ctx, args)
    spec = spectrometer(fxlib, ctx, args)
    fe = None if args.no_front_end else front_end(fxlib, ctx, args)
    chains = None if args.no_tasks else task_chains(args)
    app = None if args.no_front_end else spectrometer_app(fxlib, args)

    large = None
    if not args.no_large:
        large = large_n(fxlib, ctx, args)
    batch = None
    if not args.no_batch:
        batch = batched(fxlib, ctx, args)

    # ---- roofline of the dominant kernel of the headline step ------------------
    # per launch: the X-engine runs once per chunk, and so does the direct F
    # kernel (transform + engine-layout store of a whole chunk)
    x_flop = 8.0 * P * C * chunk
    f_bytes = I * chunk * N / 4.0 + I * chunk * C * 4.0        # packed in + FP16 channels out
    dom = max(("fstage", "xengine"), key=lambda k: k16[k])
    if dom == "xengine":
        launch = k16["xengine"]
        roof = {"kernel": k16["xengine_kernel"], "bound": "mfma",
                "achieved": x_flop / (launch * 1e-3) / 1e12, "peak": PEAK["fp16"], "unit": "TFLOP/s",
                "algorithmic": f"8*P*C*chunk = {x_flop:.3e} flop per launch"}
    else:
        if not k16["fstage_kernel"].startswith("k_fstage_direct"):
            raise SystemExit("bench: the F stage of this shape is not the direct engine-layout kernel")
        launch = k16["fstage"]
        fft_tflops = 2.5 * N * np.log2(N) * I * chunk / (launch * 1e-3) / 1e12
        roof = {"kernel": k16["fstage_kernel"], "bound": "hbm",
                "achieved": f_bytes / (launch * 1e-3) / 1e9, "peak": PEAK["hbm"], "unit": "GB/s",
                "algorithmic": f"per input-spectrum N/4 B packed in + C*4 B FP16 out; {chunk} spectra x {I} "
                               f"inputs = {f_bytes:.3e} B per launch",
                "launches_timed": k16["fstage_launches"],
                # the roof that binds: VALU issue (counters) and the FFT's flop rate
                "valu": {"fft_tflops": fft_tflops, "fft_frac_of_fp32_peak": fft_tflops / PEAK["fp32"],
                         "busy": measured_valu(k16["fstage_kernel"])}}
    roof["frac"] = roof["achieved"] / roof["peak"]
    tr = measured_traffic(roof["kernel"])
    roof["traffic"] = None if tr is None else tr["bytes"]
    if tr is not None:
        roof["traffic_source"] = tr["source"]
    roof["launch_ms"] = launch

    kern = {}
    for name, (ms, kt) in paths.items():
        for k, v in kt.items():
            kern[f"{k}_{name}"] = v
    kern["xengine_f16_tflops"] = x_flop / (k16["xengine"] * 1e-3) / 1e12
    kern["xengine_f16_mfma_busy"] = measured_busy(k16["xengine_kernel"], "mfma")
    if "f32" in paths:
        kern["xengine_f32_tflops"] = x_flop / (paths["f32"][1]["xengine"] * 1e-3) / 1e12
    if "fp8" in paths:
        kern["xengine_fp8_tflops"] = x_flop / (paths["fp8"][1]["xengine"] * 1e-3) / 1e12
    kern["chunks_per_step"] = nchunks

    def path_obj(name):
        if name not in paths:
            return None
        ms = paths[name][0]
        return {"ms_per_step": ms, "value": cmac / (ms * 1e-3),
                "input_gsps": I * float(N) * T / (ms * 1e-3) / 1e9}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "complex MAC/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms16,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16",
        "data": "synthetic",
        "config": {"workload": f"mid-size FX correlator: {args.stations} stations x 2 pol, "
                               f"{N}-pt real FFT ({C} ch), {T} spectra, rho={args.rho}, FP64 "
                               f"delay/fringe model, X-engine FP16 MFMA / FP32 accumulate",
                   "inputs": I, "nfft": N, "channels": C, "spectra": T, "products": P,
                   "chunk_spectra": chunk},
        "input_gsps": input_gsps,
        "fp32_plain": path_obj("f32"),
        "fp8_scaled": path_obj("fp8"),
        "speedup_fp16_vs_fp32": None if "f32" not in paths else paths["f32"][0] / ms16,
        "kernels_ms_per_chunk": kern,
        "err_vs_fp64": err,
        "mini_correlator": mini,
        "large_n_xengine": large,
        "batched_subbands": batch,
        "visibility_products": prod,
        "unpack": unp,
        "spectrometer": spec,
        "front_end": fe,
        "task_chains": chains,
        "spectrometer_app": app,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

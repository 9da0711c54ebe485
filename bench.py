#!/usr/bin/env python3
"""Benchmark of the F-X chain on one MI355X.

Workload (the mid-size correlator): 64 stations x 2 polarizations = 128 inputs
of 2-bit samples with correlated fraction rho = 0.02 and per-station delays of
0-50 samples, 8192-point real FFT (4096 channels), FP64 delay/fringe phase
model applied after the FFT, 65536 spectra (2^29 samples per input, 16 GiB of
packed input resident in HBM), all 8256 products integrated.

One step = reset accumulators -> F (unpack, FFT, rotate) -> X over all 65536
spectra in time chunks -> dump [p][c] complex64 visibilities.  The headline
runs the X-engine on FP16 MFMA; the plain FP32 path is timed beside it.

Prints ONE JSON line.  `value` is whole-step complex MAC/s (P*C*T / step time).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cellspe-tasklib_amd"))

METRIC = ("X-engine complex MAC/s + input GS/s, FP16/FP8 MFMA vs FP32 plain at ≤1e-3 err")
PEAK_FP16_TFLOPS = 2500.0
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--stations", type=int, default=64)
    ap.add_argument("--nfft", type=int, default=8192)
    ap.add_argument("--spectra", type=int, default=65536)
    ap.add_argument("--chunk", type=int, default=4096)
    ap.add_argument("--rho", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=20240229)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="instead of the benchmark, run the plain copy kernel over MIB MiB "
                         "(HBM bandwidth check and traffic-counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the F-X chain shards by time chunk with an RCCL reduce of the visibilities, "
              "which is designed but not built)", file=sys.stderr)
        sys.exit(2)

    import fxlib

    if args.copy_probe:
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
        return

    I, N, C, T, tau, models = workload(args)
    P = I * (I + 1) // 2
    nwords = T * N // 16            # packed words per input
    chunk = min(args.chunk, T)
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

    def run_path(prec, steps, warmup):
        corr = fxlib.Correlator(ctx, I, N, chunk, prec, use_phase=True)
        corr.set_phase_model(models)

        def step():
            corr.reset()
            corr.process(dw, nwords, 0, T)
            corr.dump(dvis)

        for _ in range(warmup):
            step()
        ctx.sync()
        e0, e1 = ctx.event(), ctx.event()
        ctx.sync()
        ctx.record(e0)
        for _ in range(steps):
            step()
        ctx.record(e1)
        ctx.sync()
        ms = ctx.elapsed_ms(e0, e1) / steps
        return corr, ms

    def kernel_times(corr, prec, reps=3):
        """Average duration of one F-stage and one X-engine launch over a chunk,
        timed with HIP events on the context stream."""
        esz = {fxlib.FX_PREC_F32: 8, fxlib.FX_PREC_F16: 4}[prec]
        dchan = ctx.alloc(C * chunk * I * esz)
        e = [ctx.event() for _ in range(4)]
        tf = tx = 0.0
        for r in range(reps):
            ctx.record(e[0])
            corr.fstage(dw, nwords, r * chunk % T, chunk, dchan, prec)
            ctx.record(e[1])
            corr.xengine(dchan, chunk)
            ctx.record(e[2])
            ctx.sync()
            tf += ctx.elapsed_ms(e[0], e[1])
            tx += ctx.elapsed_ms(e[1], e[2])
        dchan.free()
        corr.reset()
        return tf / reps, tx / reps

    result = {}
    # ---- headline: FP16 MFMA X-engine --------------------------------------
    corr16, ms16 = run_path(fxlib.FX_PREC_F16, args.steps, args.warmup)
    log(f"[bench] fp16 path: {ms16:.2f} ms/step")
    tf16, tx16 = kernel_times(corr16, fxlib.FX_PREC_F16)
    log(f"[bench] fp16 per-chunk launches: fstage {tf16:.2f} ms, xengine {tx16:.2f} ms")
    corr16.close()

    # ---- plain FP32 path ------------------------------------------------------
    ms32 = tf32 = tx32 = None
    if not args.no_fp32:
        corr32, ms32 = run_path(fxlib.FX_PREC_F32, max(1, args.steps - 1), 1)
        log(f"[bench] fp32 path: {ms32:.2f} ms/step")
        tf32, tx32 = kernel_times(corr32, fxlib.FX_PREC_F32)
        log(f"[bench] fp32 per-chunk launches: fstage {tf32:.2f} ms, xengine {tx32:.2f} ms")
        corr32.close()

    cmac = float(P) * C * T
    value = cmac / (ms16 * 1e-3)
    input_gsps = I * float(N) * T / (ms16 * 1e-3) / 1e9

    # ---- error vs the FP64 oracle on a bounded sub-run ------------------------
    err = None
    if not args.no_check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        te = 16
        streams = [dw.download_at(i * nwords * 4, np.uint32, te * N // 16) for i in range(I)]
        errs = {}
        for name, prec in (("f16", fxlib.FX_PREC_F16), ("f32", fxlib.FX_PREC_F32)):
            c = fxlib.Correlator(ctx, I, N, te, prec, use_phase=True)
            c.set_phase_model(models)
            c.process(dw, nwords, 0, te)
            errs[name] = c.visibilities()
            c.close()
        nth = min(16, os.cpu_count() or 8)
        t0 = time.time()
        ref64 = O.corr_f64(streams, N, te, models, nthreads=nth)
        log(f"[bench] FP64 oracle on {te} spectra in {time.time() - t0:.1f}s")
        err = {k: nrel(v, ref64) for k, v in errs.items()}
        err["sample"] = f"first {te} spectra of all {I} inputs, all {P} products x {C} channels"
        log(f"[bench] normwise error vs FP64 oracle: f16 {err['f16']:.2e}, f32 {err['f32']:.2e}")
        del ref64, errs

    # ---- CPU baseline (oracle FP32 generic-C chain, all host threads) --------
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        nth = min(16, os.cpu_count() or 8)
        ts = max(2, nth)
        while True:
            streams = [dw.download_at(i * nwords * 4, np.uint32, ts * N // 16) for i in range(I)]
            t0 = time.time()
            O.corr_f32(streams, N, ts, models, nthreads=nth)
            dt = time.time() - t0
            if dt > args.cpu_seconds / 3 or ts * 2 > T:
                break
            ts *= 2
        cpu = {"value": P * C * ts / dt, "unit": "complex MAC/s", "cores": nth, "kind": "port",
               "input_gsps": I * N * ts / dt / 1e9,
               "sample": f"{ts} spectra x {I} inputs x {C} channels (FP32 generic-C restatement, "
                         f"project radix-2 FFT, not FFTW), {dt:.1f}s"}
        log(f"[bench] CPU baseline: {cpu['value']:.3e} CMAC/s on {nth} threads ({ts} spectra, {dt:.1f}s)")

    # ---- roofline of the dominant kernel -------------------------------------
    nchunks = (T + chunk - 1) // chunk
    x_flop = 8.0 * P * C * chunk
    f_bytes = I * chunk * N / 4.0 + I * chunk * C * 4.0   # packed in + FP16 channelized out
    if tx16 >= tf16:
        roof = {"kernel": "k_xengine_f16", "bound": "mfma", "achieved": x_flop / (tx16 * 1e-3) / 1e12,
                "peak": PEAK_FP16_TFLOPS, "unit": "TFLOP/s"}
    else:
        roof = {"kernel": "k_fstage<4096,__half2>", "bound": "hbm",
                "achieved": f_bytes / (tf16 * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = None
    roof["launch_ms"] = tx16 if roof["bound"] == "mfma" else tf16

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
        "fp32_plain": None if ms32 is None else {
            "ms_per_step": ms32, "value": cmac / (ms32 * 1e-3),
            "input_gsps": I * float(N) * T / (ms32 * 1e-3) / 1e9},
        "speedup_fp16_vs_fp32": None if ms32 is None else ms32 / ms16,
        "kernels_ms_per_chunk": {"fstage_f16": tf16, "xengine_f16": tx16,
                                 "fstage_f32": tf32, "xengine_f32": tx32,
                                 "xengine_f16_tflops": x_flop / (tx16 * 1e-3) / 1e12,
                                 "xengine_f32_tflops": None if tx32 is None else x_flop / (tx32 * 1e-3) / 1e12,
                                 "chunks_per_step": nchunks},
        "err_vs_fp64": err,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

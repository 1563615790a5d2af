#!/usr/bin/env python3
"""Throughput of the MI355X FX correlator on one GPU.

Workload (default): 64 antennas x 2 polarisations, 2-bit real samples at
128 Msps, 4096 channels (8192-point FFT), 10 ms subintegrations of 156 FFTs,
FP16 spectra into the MFMA X-engine.  One step = the full hot path over one
batch of `--xbatch` subintegrations already resident in HBM: FP64 delay
table and fused unpack/fringe/FFT/fractional-delay F-stage (one launch of
each per `--fbatch` consecutive subints, 16 by default, the FFT indices of
the subints one contiguous run), valid-FFT counts, and one X-engine launch
accumulating the batch into the visibilities.  The data are synthetic
(seeded 2-bit Gaussian noise with a 0.1-correlated common signal): 16
consecutive subints of distinct samples (658 MB at the default workload,
more than the 256 MB Infinity Cache), cycled through with advancing global
FFT indices, so every subint of a step streams fresh input from HBM.

Prints one JSON line.  `value` is input Gsamples/s (antennas x pols x
samples processed per second) with plain launches timed by HIP events on the
engine's stream.  The same line carries the X-engine TFLOPS, the same steps
with one delay and one F-stage launch per subint, the same steps as graph
launches, the plain FP32 GPU pipeline on the same workload, the roofline of
the dominant kernel (HIP-event timings inside the timed region) and a
bounded CPU baseline (the oracle's FP32 restatement of the reference path) on
all usable host threads and on one.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from difx_tests_amd import FxEngine, workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"f16": 2500.0, "f8": 5000.0, "f32": 157.3, "f32v": 157.3}
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every 4
# cycles at 2.4 GHz (no packed FP32)
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4
OUT_BYTES_PER_SAMPLE = {"f32": 4.0, "f16": 2.0, "f8": 1.0, "f32v": 4.0}   # nchan complex bins per F real samples


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3-64ant-4096ch", choices=sorted(workloads.WORKLOADS))
    ap.add_argument("--prec", default=None, choices=["f32", "f16", "f8", "f32v"])
    ap.add_argument("--xbatch", type=int, default=0, help="subints per X-engine launch (0: the autotuner's choice)")
    ap.add_argument("--fbatch", type=int, default=0,
                    help="subints per delay / F-stage launch (0: 16, or fewer where the X batch needs it)")
    ap.add_argument("--no-autotune", action="store_true", help="default radix plan and X-engine tile, xbatch 8")
    ap.add_argument("--plan", type=int, default=None, help="F-stage radix plan (skips the autotuner)")
    ap.add_argument("--xtile", type=int, default=None, help="X-engine tile 1, 2 or 4 (skips the autotuner)")
    ap.add_argument("--no-fp32-compare", action="store_true")
    ap.add_argument("--no-subbands", action="store_true", help="skip the eight-subband batch measurement")
    ap.add_argument("--no-stream", action="store_true", help="skip the host-resident (PCIe-inclusive) measurement")
    ap.add_argument("--no-graph", action="store_true", help="skip the graph-launch measurements")
    ap.add_argument("--no-per-subint", action="store_true",
                    help="skip the measurement with one delay and one F-stage launch per subint")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="target CPU baseline work (s), split between the all-threads and one-thread legs; 0 disables")
    ap.add_argument("--seed", type=int, default=20261205)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run three plain 512 MiB device copies (calibrates the HBM traffic counters)")
    ap.add_argument("--subbands-only", action="store_true",
                    help="only run the eight-subband batch measurement (autotuned unless --no-autotune) and print it")
    ap.add_argument("--subband-modes", default="launch,graph",
                    help="modes of the subband batch measurement: launch, graph or both (comma-separated)")
    return ap.parse_args()


def copy_probe():
    """Three device-to-device copies of 512 MiB: FETCH_SIZE / WRITE_SIZE of
    this kernel against the bytes it must move give the counter scale."""
    import torch
    n = 512 * 1024 * 1024
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    a.fill_(1)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n, "copies": 3}), flush=True)


# Subints per X-engine launch the autotuner may choose from (its FFT capacity).
# Each X-engine launch reads and writes the accumulators once, and the
# F-stages right after it run slower (the clock recovers from the MFMA burst
# over a few of them), so fewer, deeper launches raise the step rate.
MAX_XBATCH = 64
NBUF = 16           # subints of distinct input held in HBM and cycled through


def fstage_batch(xbatch, want=0):
    """Subints per delay / F-stage launch: `want`, or the largest power of two
    up to NBUF that divides the X batch (and NBUF)."""
    if want:
        if xbatch % want or NBUF % want:
            raise SystemExit(f"--fbatch {want} must divide the X batch ({xbatch}) and {NBUF}")
        return want
    fb = 1
    while fb * 2 <= NBUF and xbatch % (fb * 2) == 0:
        fb *= 2
    return fb


def run_pipeline(w, prec, data, delays, fileoff, nf, steps, warmup, xbatch, tune=True, plan=None, xtile=None,
                 graph=False, timing=True, fbatch=0):
    """Returns (seconds per step, kernel ms per step dict, launch counts, slots
    per F-stage launch, tuning dict).  The input buffer holds NBUF consecutive
    subints of distinct data; subint u of the run reads subint u mod NBUF of it
    (global FFT indices (u mod NBUF)*nf + k, buffer offsets of the whole
    buffer as the reference's split does), so a step streams fresh samples.
    One delay / F-stage launch covers fbatch consecutive subints of the
    buffer, whose FFT indices form one contiguous run (fbatch 1: one launch
    per subint, each subint's slots padded to the F-stage granule).  graph:
    every F-stage launch is one recorded graph launch (fx_subint) with the
    X-engine in the step's last one; otherwise plain launches with HIP-event
    kernel timing.  xbatch 0 takes the autotuner's choice."""
    cap = (xbatch or MAX_XBATCH) * (nf + 64)
    e = FxEngine(w.nant, w.nchan, w.lo, w.bw, prec=prec, max_ffts=cap, max_bytes_per_ant=data.shape[1])
    try:
        e.set_model(delays, fileoff)
        e.upload(data)
        tuned = e.autotune(nf, reps=3) if tune else {}
        if plan is not None:
            e.set_plan(plan)
        if xtile is not None:
            e.set_xtile(xtile)
        tuned["plan_used"], tuned["xtile_used"] = e.plan(), e.xtile()
        if not xbatch:
            xbatch = tuned.get("xbatch", 8)
        tuned["xbatch_used"] = xbatch
        fb = fstage_batch(xbatch, fbatch)
        tuned["fbatch_used"] = fb
        gran = e.granule
        slots = (fb * nf + gran - 1) // gran * gran
        nblk = NBUF // fb                          # blocks of fb subints in the buffer
        its = [(r * fb * nf + np.arange(fb * nf)).astype(np.int32) for r in range(nblk)]
        nlaunch = xbatch // fb
        u = [0]

        def step():
            for s in range(nlaunch):
                it = its[u[0] % nblk]
                u[0] += 1
                if graph:
                    e.subint(it, split=nblk, fft0=s * slots, xslots=nlaunch * slots if s == nlaunch - 1 else 0)
                else:
                    e.fstage(it, nblk, s * slots)
            if not graph:
                e.xstage(0, nlaunch * slots)

        e.reset()
        for _ in range(warmup):
            step()
        e.sync()
        e.set_timing(timing)
        e.reset_timing()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        e.sync()
        t1 = time.perf_counter()
        ms, nl = e.timing()
        gms, ng = e.graph_timing()
        e.set_timing(False)
        per_step = {k: v / steps for k, v in ms.items()}
        if graph and timing:
            per_step = {"graph": gms / steps}
        return (t1 - t0) / steps, per_step, nl, slots, tuned
    finally:
        e.close()


def subband_batch(steps, warmup, seed, prec="f32", tune=True, modes=("launch", "graph")):
    """Eight uneven subbands of a 16-antenna array, one 10 ms subint per step:
    one hipGraph launch per subint (subbands as concurrent branches) against
    plain launches one subband after another.  Synthetic random 2-bit input."""
    from difx_tests_amd import FxBatch, synth
    subs = workloads.C4
    engines, iters = [], []
    for i, w in enumerate(subs):
        delays, fileoff = workloads.delay_model(w, seed + i)
        data = synth.random_bytes(w.nant, w.subint_bytes, seed + i)
        e = FxEngine(w.nant, w.nchan, w.lo, w.bw, prec=prec, max_ffts=w.nffts, max_bytes_per_ant=data.shape[1])
        e.set_model(delays, fileoff)
        e.upload(data)
        e.reset()
        engines.append(e)
        iters.append(np.arange(w.nffts, dtype=np.int32))
    samples = sum(w.samples_per_subint for w in subs)
    out = {"subbands": [{"nchan": w.nchan, "antennas": w.nant, "fs_msps": w.fs / 1e6, "ffts": w.nffts}
                        for w in subs],
           "xengine_prec": prec, "samples_per_subint": samples}
    try:
        with FxBatch(engines) as b:
            for mode in modes:
                for _ in range(warmup):
                    b.process(iters, 1, mode)
                b.sync()
                t0 = time.perf_counter()
                for _ in range(steps):
                    b.process(iters, 1, mode)
                b.sync()
                dt = (time.perf_counter() - t0) / steps
                out[mode] = {"ms_per_subint": round(dt * 1e3, 4), "gsamples_s": round(samples / dt / 1e9, 3)}
            if "launch" in out and "graph" in out:
                out["graph_speedup"] = round(out["launch"]["ms_per_subint"] / out["graph"]["ms_per_subint"], 3)
            if tune:
                choices = []
                for e, w in zip(engines, subs):
                    r = e.autotune(w.nffts, reps=3)
                    choices.append({"plan": r["plan"], "xtile": r["xtile"]})
                    e.reset()
                for _ in range(warmup):
                    b.process(iters, 1, "graph")
                b.sync()
                t0 = time.perf_counter()
                for _ in range(steps):
                    b.process(iters, 1, "graph")
                b.sync()
                dt = (time.perf_counter() - t0) / steps
                out["graph_autotuned"] = {"ms_per_subint": round(dt * 1e3, 4),
                                          "gsamples_s": round(samples / dt / 1e9, 3), "choices": choices}
            out["graph_captures"] = b.captures
    finally:
        for e in engines:
            e.close()
    return out


def streaming_ingest(w, prec, data, delays, fileoff, nf, tuned, xbatch, nsub):
    """Host-resident input: every subint's packed bytes cross PCIe into one of
    two device slots (pinned staging, copy stream) while the previous subint is
    correlated.  The staging buffers are filled once (a real caller reads its
    files into them); global FFT indices and buffer origins advance per subint.
    Returns the PCIe-inclusive rate and the same loop with copies and compute
    strictly alternating (no overlap)."""
    nbytes = data.shape[1]
    out = {"subints": nsub, "bytes_per_subint": int(nbytes) * w.nant}
    e = FxEngine(w.nant, w.nchan, w.lo, w.bw, prec=prec, max_ffts=xbatch * (nf + 64), max_bytes_per_ant=nbytes)
    try:
        e.set_model(delays, fileoff)
        e.set_plan(tuned.get("plan_used", 0))
        e.set_xtile(tuned.get("xtile_used", 0))
        for k in (0, 1):
            e.ingest_host(k)[:, :nbytes] = data
        gran = e.granule
        slots = (nf + gran - 1) // gran * gran
        for mode in ("overlapped", "serial"):
            e.reset()
            e.sync()
            t0 = time.perf_counter()
            for s in range(nsub):
                k = s % 2
                e.ingest_wait(k)
                e.ingest_submit(k, nbytes)
                if mode == "serial":
                    e.ingest_wait(k)
                    e.sync()
                b = s % xbatch
                e.fstage_ex((s * nf + np.arange(nf)).astype(np.int32), slot=k, origin=s * nf, fft0=b * slots)
                if b == xbatch - 1 or s == nsub - 1:
                    e.xstage(0, (b + 1) * slots)
                if mode == "serial":
                    e.sync()
            e.sync()
            dt = (time.perf_counter() - t0) / nsub
            out[mode] = {"ms_per_subint": round(dt * 1e3, 4),
                         "gsamples_s": round(w.samples_per_subint / dt / 1e9, 3),
                         "h2d_gb_s_if_copy_bound": round(nbytes * w.nant / dt / 1e9, 2)}
    finally:
        e.close()
    return out


def measured_traffic(workload, prec, xbatch, fbatch, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (copy-kernel calibrated), or None.
    An F-stage launch covers fbatch subints whatever the number of subints per
    X-engine launch, so for it a summary taken at any xbatch and the same
    fbatch serves."""
    pdir = os.path.join(ROOT, "profiles")
    paths = [os.path.join(pdir, f"traffic_{workload}_{prec}_xbatch{xbatch}_fbatch{fbatch}.md")]
    if kernel == "fx_fstage_kernel" and os.path.isdir(pdir):
        paths += sorted(os.path.join(pdir, n) for n in os.listdir(pdir)
                        if n.startswith(f"traffic_{workload}_{prec}_xbatch") and n.endswith(f"_fbatch{fbatch}.md"))
    for path in paths:
        if not os.path.exists(path):
            continue
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                    return int(parts[2]), os.path.relpath(path, ROOT)
    return None, None


def measured_valu(workload, prec, fbatch, kernel):
    """VALU instructions (wave64) per launch of `kernel` (fbatch subints) from
    the committed rocprofv3 SQ_INSTS_VALU summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", f"sq_counters_{workload}_{prec}_fbatch{fbatch}.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "valu_insts_per_launch" and parts[1] == kernel:
                return int(parts[2]), os.path.relpath(path, ROOT)
    return None, None


def _cpu_leg(w, data, delays, fileoff, budget_s, nthreads):
    """The oracle's FP32 restatement of the reference path on a bounded FFT
    sample, with nthreads host threads (F-stage over (FFT, antenna) pairs,
    X-engine over baselines)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as orc
    p = orc.plan(w.nchan, w.lo, w.bw)
    nbuf_ffts = (data.shape[1] * 2) // w.F - 1
    # probe with four FFTs of every antenna (per-call costs such as clearing
    # the visibilities included), then size the sample to the budget
    n0 = min(4, w.nffts)
    t0 = time.perf_counter()
    chan, valid = orc.fstage_f32(p, data, delays, fileoff, np.arange(n0, dtype=np.int32), split=nbuf_ffts // n0,
                                 nthreads=nthreads)
    orc.xstage_f32(chan, valid, nthreads=nthreads)
    t1 = time.perf_counter()
    nsample = int(max(1, min(w.nffts, n0 * budget_s / max(t1 - t0, 1e-6))))
    it = np.arange(nsample, dtype=np.int32)
    t0 = time.perf_counter()
    chan, valid = orc.fstage_f32(p, data, delays, fileoff, it, split=nbuf_ffts // nsample, nthreads=nthreads)
    t1 = time.perf_counter()
    orc.xstage_f32(chan, valid, nthreads=nthreads)
    t2 = time.perf_counter()
    samples = w.nant * 2 * nsample * w.F
    return samples / (t2 - t0) / 1e9, nsample, t1 - t0, t2 - t1


def cpu_baseline(w, data, delays, fileoff, budget_s):
    """The oracle's FP32 restatement (no IPP) timed on this host: all the
    threads this process may use (capped at 16, the CPU share of one GPU on
    the measurement boxes) and one thread."""
    visible = os.cpu_count() or 1
    usable = len(os.sched_getaffinity(0))
    nth = max(1, min(16, usable))
    v_all, n_all, f_all, x_all = _cpu_leg(w, data, delays, fileoff, budget_s / 2, nth)
    v_one, n_one, f_one, x_one = _cpu_leg(w, data, delays, fileoff, budget_s / 2, 1)
    return {
        "value": round(v_all, 6),
        "unit": "Gsamples/s",
        "cores": nth,
        "kind": "port",
        "sample": f"{n_all} FFTs x {w.nant} antennas x 2 pols (F {f_all:.2f} s + X {x_all:.2f} s on {nth} threads; "
                  f"oracle FP32 restatement of kernel.c:709-854, no IPP; host shows {visible} CPUs, "
                  f"{usable} usable by this process)",
        "single_thread": {"value": round(v_one, 6), "cores": 1,
                          "sample": f"{n_one} FFTs x {w.nant} antennas x 2 pols (F {f_one:.2f} s + X {x_one:.2f} s)"},
        "host_cpus_visible": visible,
        "host_cpus_usable": usable,
    }


def graph_vs_launch(steps, warmup, seed):
    """4 antennas x 1024 channels FP32 (the reference's CPU-runnable case, where
    launch gaps weigh most): plain launches against one graph launch per subint,
    same tuned plan, tile and subints per X-engine launch."""
    w = workloads.C1
    data, delays, fileoff = workloads.make_stream_subints(w, seed, NBUF)
    nf = w.nffts
    out = {"workload": w.name, "subints_per_step": None}
    sec_l, _, _, _, tuned = run_pipeline(w, "f32", data, delays, fileoff, nf, steps, warmup, 0, True, timing=False,
                                         fbatch=1)
    xb = tuned["xbatch_used"]
    sec_g, _, _, _, _ = run_pipeline(w, "f32", data, delays, fileoff, nf, steps, warmup, xb, False,
                                     tuned["plan_used"], tuned["xtile_used"], graph=True, timing=False, fbatch=1)
    out["subints_per_step"] = xb
    sps = xb * w.samples_per_subint
    out["launch"] = {"ms_per_step": round(sec_l * 1e3, 4), "gsamples_s": round(sps / sec_l / 1e9, 3)}
    out["graph"] = {"ms_per_step": round(sec_g * 1e3, 4), "gsamples_s": round(sps / sec_g / 1e9, 3)}
    out["graph_speedup"] = round(sec_l / sec_g, 3)
    return out


def main():
    a = parse()
    if a.copy_probe:
        copy_probe()
        return
    if a.subbands_only:
        modes = tuple(m for m in a.subband_modes.split(",") if m)
        if not modes or any(m not in ("launch", "graph") for m in modes):
            sys.exit("--subband-modes takes launch and/or graph")
        print(json.dumps(subband_batch(max(5, a.steps), 2, a.seed, tune=not a.no_autotune, modes=modes)), flush=True)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not implemented)"}))
        sys.exit(2)
    w = workloads.WORKLOADS[a.workload]
    prec = a.prec or w.prec
    data, delays, fileoff = workloads.make_stream_subints(w, a.seed, NBUF)
    nf = w.nffts
    tune = not (a.no_autotune or a.plan is not None or a.xtile is not None)
    xb_req = a.xbatch or (0 if tune else 8)
    sec, kms, nl, slots, tuned = run_pipeline(w, prec, data, delays, fileoff, nf, a.steps, a.warmup, xb_req, tune,
                                              a.plan, a.xtile, fbatch=a.fbatch)
    xbatch, fbatch = tuned["xbatch_used"], tuned["fbatch_used"]
    samples_per_step = xbatch * w.samples_per_subint
    value = samples_per_step / sec / 1e9

    # algorithmic work per launch (an F-stage launch covers fbatch subints)
    f_launch_bytes = fbatch * w.samples_per_subint * (0.25 + OUT_BYTES_PER_SAMPLE[prec])
    x_flops = 8.0 * 4 * w.nbl_with_autos * w.nchan * nf * xbatch
    f_ms = kms["fstage"] * fbatch / xbatch   # per F launch
    x_ms = kms["xstage"]                      # one X launch per step
    f_gbs = f_launch_bytes / (f_ms * 1e-3) / 1e9
    x_tflops = x_flops / (x_ms * 1e-3) / 1e12
    if kms["fstage"] >= kms["xstage"]:
        traffic, src = measured_traffic(w.name, prec, xbatch, fbatch, "fx_fstage_kernel")
        roof = {"kernel": "fx_fstage_kernel", "bound": "hbm", "achieved": round(f_gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(f_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes": int(f_launch_bytes), "avg_launch_ms": round(f_ms, 4)}
    else:
        traffic, src = measured_traffic(w.name, prec, xbatch, fbatch, "fx_xengine_kernel")
        pk = MFMA_PEAK_TFLOPS[prec]
        roof = {"kernel": "fx_xengine_kernel", "bound": "mfma", "achieved": round(x_tflops, 2), "peak": pk,
                "unit": "TFLOP/s", "frac": round(x_tflops / pk, 4), "traffic": traffic,
                "algorithmic_flops": int(x_flops), "avg_launch_ms": round(x_ms, 4)}
    if src:
        roof["traffic_source"] = src
    # the F-stage's real limiter: VALU issue (non-packed FP32), its profiled
    # instruction count over the live launch time against the issue peak
    valu_roof = None
    ninst, vsrc = measured_valu(w.name, prec, fbatch, "fx_fstage_kernel")
    if ninst:
        g = ninst / (f_ms * 1e-3) / 1e9
        valu_roof = {"kernel": "fx_fstage_kernel", "bound": "valu", "achieved": round(g, 1),
                     "peak": VALU_PEAK_GINST, "unit": "G wave64-instructions/s",
                     "frac": round(g / VALU_PEAK_GINST, 4), "insts_per_launch": ninst, "source": vsrc}

    out = {
        "metric": "input Gsamples/s/GPU + X-engine TFLOPS; FP16-MFMA vs FP32 at <=1e-3 rel err",
        "value": round(value, 3),
        "unit": "Gsamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec,
        "data": "synthetic",
        "config": {"workload": w.name, "antennas": w.nant, "pols": 2, "nchan": w.nchan, "fs_msps": w.fs / 1e6,
                   "subint_ms": w.subint_s * 1e3, "ffts_per_subint": nf, "subints_per_step": xbatch,
                   "subints_per_fstage_launch": fbatch,
                   "xengine_prec": prec, "parallelism": "single GPU",
                   "input_in_hbm_bytes": int(data.nbytes), "distinct_subints_cycled": NBUF},
        "xengine_tflops": round(x_tflops, 2),
        "fstage_gbs": round(f_gbs, 1),
        "kernel_ms_per_step": {k: round(v, 4) for k, v in kms.items()},
        "realtime_factor": round((xbatch * w.subint_s) / sec, 2),
        "tuning": tuned,
        "roofline": roof,
        "valu_roofline": valu_roof,
    }
    if not a.no_fp32_compare and prec not in ("f32", "f32v"):
        # the plain FP32 GPU pipeline (vector-ALU X-engine) and the FP32 MFMA one
        for key, p32 in (("fp32_pipeline", "f32v"), ("fp32_mfma_pipeline", "f32")):
            s32, k32, _, _, t32 = run_pipeline(w, p32, data, delays, fileoff, nf, max(3, a.steps // 2), 1, xbatch,
                                               tune, fbatch=fbatch)
            v32 = samples_per_step / s32 / 1e9
            out[key] = {"xengine": p32, "value": round(v32, 3), "ms_per_step": round(s32 * 1e3, 4),
                        "kernel_ms_per_step": {k: round(v, 4) for k, v in k32.items()},
                        "tuning": t32, "speedup_vs_this": round(value / v32, 3)}
    if fbatch > 1 and not a.no_per_subint:
        # the same steps with one delay and one F-stage launch per subint
        s1, k1, _, _, _ = run_pipeline(w, prec, data, delays, fileoff, nf, a.steps, a.warmup, xbatch, False,
                                       tuned["plan_used"], tuned["xtile_used"], fbatch=1)
        out["per_subint_launches"] = {"value": round(samples_per_step / s1 / 1e9, 3),
                                      "ms_per_step": round(s1 * 1e3, 4),
                                      "kernel_ms_per_step": {k: round(v, 4) for k, v in k1.items()},
                                      "batched_speedup": round(s1 / sec, 3)}
    if not a.no_graph:
        # the same steps with every F-stage launch (and its delay kernel) as
        # one recorded graph launch, one per subint and one per fbatch subints
        for key, fbg in (("graph_launch", 1), ("graph_launch_batched", fbatch)):
            if key == "graph_launch_batched" and fbatch == 1:
                continue
            sg, _, _, _, _ = run_pipeline(w, prec, data, delays, fileoff, nf, a.steps, a.warmup, xbatch, False,
                                          tuned["plan_used"], tuned["xtile_used"], graph=True, timing=False,
                                          fbatch=fbg)
            out[key] = {"value": round(samples_per_step / sg / 1e9, 3), "ms_per_step": round(sg * 1e3, 4),
                        "subints_per_graph_launch": fbg, "vs_plain_launches": round(sec / sg, 3)}
        out["c1_graph_vs_launch"] = graph_vs_launch(max(10, a.steps), a.warmup, a.seed + 1)
    if not a.no_stream:
        sub = np.ascontiguousarray(data[:, : (nf + 1) * w.F // 2])
        out["pcie_inclusive"] = streaming_ingest(w, prec, sub, delays, fileoff, nf, tuned, xbatch,
                                                 max(16, 2 * xbatch))
    if not a.no_subbands:
        out["subband_batch"] = subband_batch(max(5, a.steps), 2, a.seed, tune=tune)
    if a.cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(w, data, delays, fileoff, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

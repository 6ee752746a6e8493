#!/usr/bin/env python3
"""One-GPU benchmark of the RL / OMW deconvolution hot path.

Default workload (the headline metric): decon_lucy_omw_function with 2
iterations on a 512x512x256 uint16 volume (20k Gaussian blobs, smooth
background, blurred by a 64x64x101 PSF sheared 31.5 deg, Poisson noise),
Background = 100, Wiener-Butterworth back projector (alpha 0.005, order 10),
OTFs cached on the device, input and uint16 output resident in HBM.  One step
= one decon call.

Prints ONE JSON line.  `value` is deconvolved Mvoxel/s on one GPU; `roofline`
is for the pass with the largest share of device time (algorithmic bytes per
launch / mean launch time, HIP events on the engine's stream); `plain_gpu` is
the same call on the MATLAB-structured unfused path with OTFs rebuilt per
call; `cpu_baseline` times the oracle (the CPU single-precision restatement)
on the host cores for the same call.  The same line carries the metric's
other halves: `batched` (64 random-size chunks, OMW-2 and Biggs-10 in
batched launches, Mvoxel/s against the chunks' HBM roofline and against the
plain path) and `served` (the continuous-batching server under Poisson
arrivals at 0.7x its measured capacity for --duration seconds: p50 / p99
latency, with the plain path's on the same arrival trace).

Other workloads (--workload): rl128 (Biggs RL 15 it. on 128x128x64), batched
(64 random-size chunks, OMW-2 and RL-10 in batched launches), served
(continuous-batching server, Poisson arrivals, p50/p99 vs the plain path),
large (2048x2048x512, OMW-2 and Biggs-10), deskew (deskewFrame3D of a
512x512x256 decon result), rotate (rotateFrame3D of the deskewed volume),
probes (the passes' access patterns as plain copies: HBM ceilings).
"""
import argparse
import ctypes
import json
import math
import os
import sys
import threading
import time

import numpy as np

# Hardware queues per process, read when the HIP runtime initialises: batched
# and served steps run up to eight chunk-set streams (four by default) beside
# the engine's and the copy streams, and HIP's default of 4 queues leaves them
# two.  Raised to 12 when lower (a setting above 12 is kept).
try:
    _q = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
except ValueError:
    _q = 4
if _q < 12:
    os.environ["GPU_MAX_HW_QUEUES"] = "12"

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E
PLAIN = 3  # PK_FLAG_UNFUSED | PK_FLAG_NO_OTF_CACHE


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="omw", choices=["omw", "rl128", "batched", "served", "large", "copy",
                                                                  "deskew", "rotate", "probes"])
    ap.add_argument("--shape", type=int, nargs=3, default=None)
    ap.add_argument("--iters", type=int, default=None)
    ap.add_argument("--duration", type=float, default=60.0, help="served: seconds of Poisson arrivals")
    ap.add_argument("--max-active", type=int, default=32, help="served: requests in flight at most")
    ap.add_argument("--plain-seconds", type=float, default=15.0,
                    help="served: seconds of the same arrival trace replayed on the plain path")
    ap.add_argument("--no-batched", action="store_true", help="default workload: skip the batched section (64 chunks)")
    ap.add_argument("--no-served", action="store_true", help="default workload: skip the served section (request mix)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the parity check against the oracle")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-path comparison")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch shapes instead of measuring them for the workload's sizes")
    ap.add_argument("--bp", default="wb", choices=["wb", "omw"],
                    help="OMW back projector: Wiener-Butterworth (BackProjector) or omw_backprojector_generation")
    ap.add_argument("--events-in-timed", action="store_true",
                    help="record the per-pass HIP events inside the timed K steps (the iteration loops then run "
                         "kernel by kernel instead of as HIP graphs); default: time the K steps as they run in "
                         "production and take the per-pass durations from a second, instrumented K-step run")
    return ap.parse_args()


def col_major_cuda(a, torch):
    return torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).cuda().permute(2, 1, 0)


def empty_cm(shape, dtype, torch):
    return torch.empty((shape[2], shape[1], shape[0]), dtype=dtype, device="cuda").permute(2, 1, 0)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pk_oracle

    pk_oracle.build()
    return pk_oracle


BP_DESC = {"wb": "BackProjector(psf, 'wiener-butterworth', 0.005, 1, 10): Wiener-Butterworth, alpha 0.005, "
                 "order 10, cutoff gain and resolution by the reference defaults",
           "omw": "omw_backprojector_generation(psf, 0.005, [], OTFCumThresh 0.9, hannWinBounds [0.8 1])"}


def back_projector(psf, kind="wb"):
    """The back projector PSF_b (single) for OMW decon: Wiener-Butterworth
    (BackProjector, alpha 0.005, order 10) or the OMW generator."""
    from petakit5d_amd import BackProjector, omw_backprojector_generation

    t = time.perf_counter()
    if kind == "wb":
        b = BackProjector(psf, "wiener-butterworth", 0.005, 1, 10)[0].astype(np.float32)
    else:
        b, _, _, _ = omw_backprojector_generation(psf, 0.005, None, OTFCumThresh=0.9, hannWinBounds=[0.8, 1])
    return b, time.perf_counter() - t


class Caller:
    """Direct C-ABI calls on device-resident volumes (no per-call Python copies)."""

    def __init__(self, ctx, L, method, I, psf_f, psf_b, iters, background=100.0, flags=0, out_dtype=None):
        import torch

        from petakit5d_amd import _lib
        self.ctx, self.L, self.method, self.iters = ctx, L, method, iters
        shape = tuple(I.shape)
        self.out = empty_cm(shape, out_dtype or torch.uint16, torch)
        code_in = _lib.PK_UINT16 if I.dtype == torch.uint16 else _lib.PK_SINGLE
        self.vin = _lib.PkVol(ctypes.c_void_p(I.data_ptr()), code_in, _lib.PK_MEM_DEVICE, (ctypes.c_int64 * 3)(*shape))
        out16 = self.out.dtype == torch.uint16
        self.vout = _lib.PkVol(ctypes.c_void_p(self.out.data_ptr()), _lib.PK_UINT16 if out16 else _lib.PK_SINGLE,
                               _lib.PK_MEM_DEVICE, (ctypes.c_int64 * 3)(*shape))
        self.opts = _lib.PkOpts()
        L.pk_opts_default(ctypes.byref(self.opts))
        self.opts.background = background
        self.opts.flags = flags
        self.opts.save16bit = 1 if out16 else 0
        self.pf = np.asfortranarray(psf_f, dtype=np.float32)
        self.pb = np.asfortranarray(psf_b if psf_b is not None else psf_f, dtype=np.float32)
        self.fd, self.bd = _lib.i64x3(self.pf.shape), _lib.i64x3(self.pb.shape)

    def __call__(self):
        L, c = self.L, self.ctx
        if self.method == "omw":
            st = L.pk_decon_lucy_omw(c.handle, ctypes.byref(self.vin), ctypes.c_void_p(self.pf.ctypes.data), self.fd,
                                     ctypes.c_void_p(self.pb.ctypes.data), self.bd, self.iters, ctypes.byref(self.opts),
                                     ctypes.byref(self.vout), None)
        else:
            st = L.pk_decon_lucy(c.handle, ctypes.byref(self.vin), ctypes.c_void_p(self.pf.ctypes.data), self.fd,
                                 self.iters, ctypes.byref(self.opts), ctypes.byref(self.vout), None)
        c.check(st)


def timed(fn, steps, warmup, torch, ctx=None, events_in_timed=True):
    """W warmup calls, then K timed calls.  With a context, every pass of the
    K calls is bracketed by HIP events on the engine's stream (per-pass
    durations); events_in_timed=False times the K calls without events and
    records the per-pass durations over a second run of K calls."""
    for _ in range(max(warmup, 1)):
        fn()
    torch.cuda.synchronize()

    def loop(profile):
        if profile:
            ctx.profile(True)
        lat = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ts = time.perf_counter()
            fn()
            lat.append(time.perf_counter() - ts)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        prof = None
        if profile:
            prof = ctx.profile_read()
            ctx.profile(False)
        return (t1 - t0) / steps * 1e3, lat, prof

    if ctx is None:
        return loop(False)
    if events_in_timed:
        return loop(True)
    ms, lat, _ = loop(False)
    ms_ev, _, prof = loop(True)
    prof = dict(prof)
    prof["__ms_with_events"] = ms_ev
    return ms, lat, prof


def copy_bandwidth(torch, nbytes=1 << 30, reps=10):
    """Device-to-device copy of nbytes (read + write) with torch's copy kernel:
    the achievable HBM rate on this device, for reference."""
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2.0 * nbytes / dt / 1e9


def measured_traffic(pass_name, workload):
    """HBM bytes per dispatch of a pass, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs committed under profiles/ (calibrated against a copy of
    known size); None when no measurement exists for this workload."""
    path = os.path.join(ROOT, "profiles", f"pk_kernels_{workload}_hbm_bytes.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get(pass_name)
        return None if k is None else round(k["traffic_bytes"])
    except (OSError, ValueError, KeyError):
        return None


def roofline(prof, steps, workload=None):
    prof = {k: v for k, v in prof.items() if not k.startswith("__")}
    name, (launches, tot_ms, bpl) = max(prof.items(), key=lambda kv: kv[1][1])
    avg_ms = tot_ms / launches
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    passes = {k: {"launches_per_step": v[0] / steps, "avg_ms": round(v[1] / v[0], 4),
                  "GBps": round(v[2] / (v[1] / v[0] * 1e-3) / 1e9, 1) if v[2] else None}
              for k, v in sorted(prof.items())}
    return ({"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
             "frac": round(achieved / HBM_PEAK_GBPS, 4), "algorithmic_bytes": round(bpl),
             "traffic": measured_traffic(name, workload) if workload else None},
            passes, sum(v[1] for v in prof.values()) / steps)


def progress(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


# the headline metric; the default run's line carries it with its parts
HEADLINE_METRIC = "deconvolved Mvoxel/s (OMW, 2 it) + p99 latency at fixed load vs plain GPU path"


def base_result(metric, value, unit, args, ms, config):
    return {"metric": metric, "value": round(value, 2), "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": config}


# ---------------------------------------------------------------------------------------------- workloads
def wl_omw(args, torch, pk, L):
    from petakit5d_amd import synth
    shape = tuple(args.shape or (512, 512, 256))
    iters = args.iters or 2
    t_gen = time.perf_counter()
    I, psf = synth.blob_case(shape=shape)
    psf_b, t_bp = back_projector(psf, args.bp)
    t_gen = time.perf_counter() - t_gen
    ctx = pk.Context(0)
    It = col_major_cuda(I, torch)
    tune = None
    if not args.no_autotune:  # per-size launch-shape / pass-order measurement, before the timed region
        t = time.perf_counter()
        tune = {"report": ctx.autotune(shape), "s": round(time.perf_counter() - t, 3)}
    call = Caller(ctx, L, "omw", It, psf, psf_b, iters)
    ms, lat, prof = timed(call, args.steps, args.warmup, torch, ctx, events_in_timed=args.events_in_timed)
    nvox = float(np.prod(shape))
    value = nvox / 1e6 / (ms / 1e3)
    rl, passes, dev_ms = roofline(prof, args.steps, "omw2_512x512x256" if shape == (512, 512, 256) and iters == 2
                                  else None)
    N = nvox
    S = (shape[0] // 2 + 1) * shape[1] * shape[2]
    model = iters * (24 * N + 176 * S) + 12 * N
    res = base_result("deconvolved Mvoxel/s (OMW, 2 it)", value, "Mvoxel/s", args, ms, {
        "workload": f"OMW decon, {iters} iterations, {shape[0]}x{shape[1]}x{shape[2]} uint16, 20k blobs + sinusoidal "
                    f"background, 64x64x101 PSF sheared 31.5 deg, Poisson; {args.bp.upper()} back projector alpha 0.005; "
                    "OTFs cached, input/output resident in HBM; one step = one decon call",
        "volume": list(shape), "psf": [64, 64, 101], "iterations": iters,
        "back_projector": BP_DESC[args.bp],
        "background": 100, "output": "uint16"})
    res["roofline"] = rl
    # algorithmic bytes of the engine's passes per call (each operand of each
    # pass read or written once), against the wall-clock time per call
    eng = sum(v[0] * v[2] for k, v in prof.items() if not k.startswith("__")) / args.steps
    res["call_model"] = {"bytes_per_call": round(eng), "GBps_vs_wall": round(eng / (ms * 1e-3) / 1e9, 1),
                         "frac_vs_wall": round(eng / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
                         "device_ms_per_step": round(dev_ms, 4),
                         "full_spectrum_model_bytes": model}
    res["passes"] = passes
    res["latency_ms"] = {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                         "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)}
    if "__ms_with_events" in prof:
        res["call_model"]["ms_per_step_with_events"] = round(prof["__ms_with_events"], 4)
    res["hbm_copy_GBps"] = round(copy_bandwidth(torch), 1)
    if tune:
        res["autotune"] = tune
    res["gen_s"] = round(t_gen, 1)
    res["backprojector_s"] = round(t_bp, 2)
    if not args.no_plain:
        plain = Caller(ctx, L, "omw", It, psf, psf_b, iters, flags=PLAIN)
        pms, _, _ = timed(plain, max(2, args.steps // 3), 1, torch)
        res["plain_gpu"] = {"value": round(nvox / 1e6 / (pms / 1e3), 2), "unit": "Mvoxel/s", "ms_per_step": round(pms, 3),
                            "path": "MATLAB-structured: one kernel per gpuArray statement, OTFs rebuilt per call",
                            "speedup": round(value / (nvox / 1e6 / (pms / 1e3)), 2)}
    # PCIe-inclusive rate (host uint16 in and out through the same entry point)
    t = time.perf_counter()
    for _ in range(3):
        pk.decon_lucy_omw_function(I, psf, psf_b, iters, Background=100, ctx=ctx)
    res["host_io_Mvoxel_s"] = round(nvox / 1e6 / ((time.perf_counter() - t) / 3), 1)
    progress("cpu baseline and parity")
    cpu_and_parity(args, res, ctx, pk, I, psf, psf_b, iters, "omw", nvox, shape)
    if shape == (512, 512, 256) and iters == 2:
        # the headline metric's other halves: batched (64 chunks) and served at fixed load (request mix)
        ctx.close()
        del It, call
        torch.cuda.empty_cache()
        if not args.no_batched:
            res["batched"] = batched_section(args, torch, pk, psf, psf_b, args.steps)
        if not args.no_served:
            res["served"] = served_section(args, torch, pk)
        if not (args.no_batched or args.no_served or args.no_plain):
            # the headline metric in full: value is its Mvoxel/s part, the
            # rest lives in the objects named here
            res["metric"] = HEADLINE_METRIC
            res["metric_parts"] = {
                "deconvolved Mvoxel/s (OMW, 2 it)": "value (512x512x256, one call per step); "
                                                    "batched.omw2.Mvoxel_s (64 chunks)",
                "p99 latency at fixed load": "served.latency_ms.p99 (request mix, Poisson arrivals at 0.7x "
                                             "capacity)",
                "vs plain GPU path": "plain_gpu.speedup, batched.*.speedup_vs_plain, served.plain_gpu"}
    return res


def cpu_and_parity(args, res, ctx, pk, I, psf, psf_b, iters, method, nvox, shape):
    if args.no_check and args.no_cpu_baseline:
        return
    po = oracle()
    if method == "omw":
        Jg, _ = pk.decon_lucy_omw_function(I, psf, psf_b, iters, Background=100, save16bit=False, ctx=ctx)
    else:
        Jg, _, _ = pk.decon_lucy_function(I, psf, iters, Background=100, save16bit=False, ctx=ctx)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def cpu_call():
        if method == "omw":
            return po.decon_lucy_omw_function(I, psf, psf_b, iters, Background=100, save16bit=False)[0]
        return po.decon_lucy_function(I, psf, iters, Background=100, save16bit=False)[0]

    Jc = cpu_call()  # the parity reference, and the warm-up of the timed calls
    if not args.no_cpu_baseline:
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            cpu_call()
            times.append(time.perf_counter() - t0)
        tc = float(np.median(times))
        res["cpu_baseline"] = {"value": round(nvox / 1e6 / tc, 3), "unit": "Mvoxel/s", "cores": threads, "kind": "port",
                               "sample": f"full {shape[0]}x{shape[1]}x{shape[2]} {method.upper()}-{iters} calls "
                                         "(oracle FP32, full-spectrum FFTs, OTF build included): median of 3 after "
                                         f"a warm-up call, {tc:.1f} s each"}
    if not args.no_check:
        par = {"rel_l2_vs_cpu_fp32": rel_l2(Jg, Jc)}
        if nvox <= 2 ** 27:
            if method == "omw":
                J64, _ = po.decon_lucy_omw_function(I, psf, psf_b, iters, Background=100, save16bit=False,
                                                    precision="f64")
            else:
                J64, _, _ = po.decon_lucy_function(I, psf, iters, Background=100, save16bit=False, precision="f64")
            par["rel_l2_vs_fp64"] = rel_l2(Jg, J64)
            par["cpu_fp32_vs_fp64"] = rel_l2(Jc, J64)
        res["parity"] = par


def wl_rl128(args, torch, pk, L):
    from petakit5d_amd import synth
    I, psf = synth.emitter_case()
    iters = args.iters or 15
    ctx = pk.Context(0)
    It = col_major_cuda(I, torch)
    tune = None
    if not args.no_autotune:  # launch shapes measured for the size before the timed region (shared by the batch)
        t = time.perf_counter()
        tune = {"report": ctx.autotune(I.shape), "s": round(time.perf_counter() - t, 3)}
    call = Caller(ctx, L, "rl", It, psf, None, iters)
    ms, lat, prof = timed(call, args.steps, args.warmup, torch, ctx, events_in_timed=args.events_in_timed)
    nvox = float(I.size)
    res = base_result("deconvolved Mvoxel/s (Biggs RL, 15 it)", nvox / 1e6 / (ms / 1e3), "Mvoxel/s", args, ms, {
        "workload": "Biggs RL 15 iterations, 128x128x64 uint16, 200 emitters, 31x31x41 PSF, background 100, Poisson; "
                    "one step = one decon call", "volume": list(I.shape), "iterations": iters})
    res["roofline"], res["passes"], _ = roofline(prof, args.steps)
    if tune:
        res["autotune"] = tune
    if not args.no_plain:
        pms, _, _ = timed(Caller(ctx, L, "rl", It, psf, None, iters, flags=PLAIN), max(2, args.steps // 3), 1, torch)
        res["plain_gpu"] = {"value": round(nvox / 1e6 / (pms / 1e3), 2), "ms_per_step": round(pms, 3)}
    # A single 128x128x64 call is bound by launch latency (each pass is a few
    # microseconds of HBM work); volumes this small belong in batches: 32
    # independent volumes of the same size in one batched call
    vols = [It] + [col_major_cuda(synth.emitter_volume(I.shape, psf, 200, (500, 3000), 100.0, 100 + k), torch)
                   for k in range(31)]
    fn = lambda: pk.decon_batch(vols, psf, None, iters, method="rl", Background=100, ctx=ctx)  # noqa: E731
    bms, _, _ = timed(fn, max(2, args.steps // 5), 1, torch)
    res["batched_32"] = {"Mvoxel_s": round(32 * nvox / 1e6 / (bms / 1e3), 1), "ms_per_batch": round(bms, 3),
                         "volumes": 32, "how": "32 independent 128x128x64 volumes in one pk_decon_batch call"}
    cpu_and_parity(args, res, ctx, pk, I, psf, None, iters, "rl", nvox, I.shape)
    return res


def _chunk_sizes(n, seed):
    from petakit5d_amd import synth
    g = synth.rng(3, seed)
    return [(int(g.integers(128, 385)), int(g.integers(128, 385)), int(g.integers(64, 257))) for _ in range(n)]


def chunk_model_bytes(shape, iters, method):
    """Algorithmic bytes of one decon call on the padded size of a chunk (the
    mandated pass design: 24N + 176S per OMW iteration, 40N + 176S per Biggs
    iteration, 12N set-up; N padded voxels, S half-spectrum elements)."""
    from petakit5d_amd.decon import findGoodFactorNumber
    n1, n2, n3 = (findGoodFactorNumber(n, 1, False, maxPrime=5) for n in shape)
    N = n1 * n2 * n3
    S = (n1 // 2 + 1) * n2 * n3
    per = (24 * N + 176 * S) if method == "omw" else (40 * N + 176 * S)
    return iters * per + 12 * N


def batched_section(args, torch, pk, psf, psf_b, steps):
    """64 chunks (x, y ~ U[128, 384], z ~ U[64, 256]) cropped from one
    blob volume, OMW-2 and Biggs-10 through pk_decon_batch; Mvoxel/s, the
    fraction of the chunks' HBM roofline, one fused call per chunk and the
    plain path beside it."""
    from petakit5d_amd import synth
    sizes = _chunk_sizes(64, 0)
    # chunks are crops of one blob volume (the 512x512x256 generator at the same blob density)
    pool = synth._blob_volume_torch((384, 384, 256), psf, 11250, 3, 100.0, as_tensor=True)
    g = synth.rng(3, 1)
    chunks = []
    for s in sizes:
        o = [int(g.integers(0, 384 - s[0] + 1)), int(g.integers(0, 384 - s[1] + 1)), int(g.integers(0, 256 - s[2] + 1))]
        chunks.append(pool[o[0]:o[0] + s[0], o[1]:o[1] + s[1], o[2]:o[2] + s[2]].permute(2, 1, 0).contiguous()
                      .permute(2, 1, 0))
    del pool
    ctx = pk.Context(0)
    nvox = float(sum(np.prod(s) for s in sizes))
    from petakit5d_amd.decon import findGoodFactorNumber
    npad = float(sum(np.prod([findGoodFactorNumber(n, 1, False, maxPrime=5) for n in s]) for s in sizes))
    out = {"chunks": 64, "voxels": nvox, "padded_voxels": npad}
    for method, iters in (("omw", 2), ("rl", 10)):
        progress(f"batched {method}-{iters}: 64 chunks")
        fl = 0 if args.no_autotune else 4  # PK_FLAG_AUTOTUNE: shapes measured per chunk size in the warm-up batch
        fn = lambda: pk.decon_batch(chunks, psf, psf_b, iters, method=method, Background=100, ctx=ctx,  # noqa: E731
                                    flags=fl)
        nb = max(2, steps // 3)
        ms, _, prof = timed(fn, nb, 1, torch, ctx, events_in_timed=args.events_in_timed)
        prof.pop("__ms_with_events", None)
        mvs = nvox / 1e6 / (ms / 1e3)
        model = float(sum(chunk_model_bytes(s, iters, method) for s in sizes))
        roof = nvox / 1e6 / (model / (HBM_PEAK_GBPS * 1e9))
        rec = {"Mvoxel_s": round(mvs, 1), "ms_per_batch": round(ms, 2),
               # the same batch counted in transform-size (2/3/5-smooth padded) voxels
               "padded_Mvoxel_s": round(npad / 1e6 / (ms / 1e3), 1),
               "roofline_Mvoxel_s": round(roof, 1), "roofline_frac": round(mvs / roof, 4),
               "model_bytes_per_batch": model,
               "launches_per_batch": int(sum(v[0] for v in prof.values()) / nb),
               "passes_ms_per_batch": {k: round(v[1] / nb, 3) for k, v in sorted(prof.items())}}

        # one fused call per chunk (chunks padded by the Python wrapper when not smooth)
        def seq(plain=False):
            for c in chunks:
                pf = pk.crop_psf_to_data(psf, c.shape)
                pb = pk.crop_psf_to_data(psf_b, c.shape)
                if method == "omw":
                    pk.decon_lucy_omw_function(c, pf, pb, iters, Background=100, pad=True, ctx=ctx,
                                               flags=PLAIN if plain else 0)
                else:
                    pk.decon_lucy_function(c, pf, iters, Background=100, pad=True, ctx=ctx, flags=PLAIN if plain else 0)
        sms, _, _ = timed(seq, 1, 1, torch)
        rec["one_call_per_chunk_Mvoxel_s"] = round(nvox / 1e6 / (sms / 1e3), 1)
        if not args.no_plain:
            pms, _, _ = timed(lambda: seq(True), 1, 0, torch)
            rec["plain_gpu_Mvoxel_s"] = round(nvox / 1e6 / (pms / 1e3), 1)
            rec["speedup_vs_plain"] = round(rec["Mvoxel_s"] / rec["plain_gpu_Mvoxel_s"], 2)
        out[f"{method}{iters}"] = rec
    ctx.close()
    return out


def wl_batched(args, torch, pk, L):
    from petakit5d_amd import synth
    psf = synth.sheared_psf()
    psf_b, _ = back_projector(psf, args.bp)
    out = batched_section(args, torch, pk, psf, psf_b, args.steps)
    res = base_result("deconvolved Mvoxel/s (batched OMW, 2 it)", out["omw2"]["Mvoxel_s"], "Mvoxel/s", args,
                      out["omw2"]["ms_per_batch"], {
                          "workload": "64 chunks, x,y ~ U[128,384], z ~ U[64,256] (padded to 2/3/5-smooth), cropped "
                                      "from a blob volume, one shared 64x64x101 PSF (centre-cropped for chunks < 101 "
                                      "in z), device-resident uint16; one step = one batch", "chunks": 64,
                          "voxels": out["voxels"]})
    res["batched"] = out
    return res


def served_section(args, torch, pk, half_load=False):
    """A continuous-batching server fed Poisson arrivals at 0.7x its
    measured capacity for args.duration seconds; p50 / p99 latency, Mvoxel/s,
    and the plain path (one unfused call per request, FIFO) on the first
    args.plain_seconds of the same arrival trace."""
    from petakit5d_amd import synth
    g = synth.rng(4, 0)
    psfs = [synth.gaussian_psf((64, 64, 101), 1.0, 2.5), synth.gaussian_psf((64, 64, 101), 1.3, 3.2)]
    bps = [back_projector(p, args.bp)[0] for p in psfs]
    pool = synth._blob_volume_torch((1024, 1024, 256), psfs[0], 20000, 4, 100.0, as_tensor=True)
    nshape = 24
    shapes = []
    for _ in range(nshape):
        x = int(round(np.exp(g.uniform(np.log(96), np.log(1024)))))
        y = int(round(np.exp(g.uniform(np.log(96), np.log(1024)))))
        z = int(round(np.exp(g.uniform(np.log(64), np.log(256)))))
        shapes.append((x, y, z))
    # the client's volumes sit in pinned host memory (uint16, column-major), as a
    # serving front end receives them; the server reads them in place
    vols = []
    for s in shapes:
        t = torch.empty((s[2], s[1], s[0]), dtype=torch.uint16, pin_memory=True)
        t.copy_(pool[:s[0], :s[1], :s[2]].permute(2, 1, 0))
        vols.append(t.permute(2, 1, 0))
    del pool
    torch.cuda.empty_cache()

    def job(i):
        meth = "omw" if g.random() < 0.7 else "rl"
        ch = int(g.integers(0, 2))
        return vols[i % nshape], meth, (2 if meth == "omw" else 15), ch

    ctx = pk.Context(0, 200 << 30)
    srv = pk.Server(ctx, max_active=args.max_active)

    fl = 0 if args.no_autotune else 4  # PK_FLAG_AUTOTUNE: each size is measured the first time it arrives (warm-up)

    def submit(v, meth, it, ch):
        return srv.submit(v, psfs[ch], bps[ch], it, method=meth, Background=100, flags=fl)

    # warm plans / OTF cache for every (shape, channel), then measure saturated capacity
    warm = [submit(vols[i], "omw", 1, c) for i in range(nshape) for c in (0, 1)]
    for r in warm:
        r.result()
    del warm  # results consumed: their pinned staging returns to the host allocator's cache
    # saturation capacity: 144 requests submitted at once; the server admits
    # max_active of them and the rest join as others finish, so the device stays
    # saturated until the queue drains (with 48 requests in all, every one admitted
    # at once, the last long requests ran nearly alone)
    jobs = [job(i) for i in range(144)]
    t0 = time.perf_counter()
    reqs = [submit(*j) for j in jobs]
    for r in reqs:
        r.result()
    cap_s = time.perf_counter() - t0
    cap_vox = sum(float(np.prod(j[0].shape)) for j in jobs)
    del reqs
    cap_rps = len(jobs) / cap_s
    qps = 0.7 * cap_rps
    import queue as _queue

    def open_loop(rate, seconds):
        """Poisson arrivals at `rate` for `seconds`; a client thread consumes
        results in submission order and drops them (pinned staging and outputs
        return to the host allocator's cache).  Returns (trace, wall s,
        latencies, per-request detail)."""
        arr, t = [], 0.0
        while t < seconds:
            t += g.exponential(1.0 / rate)
            arr.append((t, job(len(arr))))
        pending = _queue.Queue()
        lat, detail = [], []

        def consumer():
            while True:
                item = pending.get()
                if item is None:
                    return
                r, sh, meth = item
                r.result()
                inf = r.info()
                lat.append(inf["latency_ms"])
                detail.append((round(inf["latency_ms"], 1), round(inf["queue_ms"], 1), meth, list(sh)))

        cth = threading.Thread(target=consumer)
        cth.start()
        t0 = time.perf_counter()
        mark = 20.0
        for at, j in arr:
            d = at - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            pending.put((submit(*j), j[0].shape, j[1]))
            if at >= mark:
                progress(f"served: {at:.0f} s of arrivals, {len(lat)} requests done")
                mark += 20.0
        pending.put(None)
        cth.join()
        return arr, time.perf_counter() - t0, lat, detail

    # open-loop Poisson arrivals at 0.7x capacity (the reported load)
    progress(f"served: capacity {cap_rps:.1f} req/s; Poisson arrivals at {qps:.1f} req/s for {args.duration:.0f} s")
    arr, wall, lat, detail = open_loop(qps, args.duration)
    vox = sum(float(np.prod(j[0].shape)) for _, j in arr)
    res = {"workload": f"Poisson arrivals at {qps:.2f}/s (0.7x measured capacity) for {args.duration:.0f} s; "
                       "sizes log-uniform 96x96x64..1024x1024x256 (24 shapes), 2 channels "
                       "(64x64x101 PSFs sigma 1.0/2.5, 1.3/3.2), 70% OMW-2 / 30% Biggs-15, HBM budget "
                       "200 GB, host uint16 in/out (client volumes in pinned memory)",
           "seconds": args.duration, "requests": len(arr), "req_per_s": round(qps, 2),
           "Mvoxel_s": round(vox / 1e6 / wall, 1)}
    res["latency_ms"] = {"p50": round(float(np.percentile(lat, 50)), 2), "p90": round(float(np.percentile(lat, 90)), 2),
                         "p99": round(float(np.percentile(lat, 99)), 2)}
    detail.sort(key=lambda d: -d[0])
    res["slowest"] = [{"latency_ms": d[0], "queue_ms": d[1], "method": d[2], "shape": d[3]} for d in detail[:3]]
    if half_load:
        arr_h, wall_h, lat_h, _ = open_loop(0.5 * cap_rps, args.duration / 2)
        vox_h = sum(float(np.prod(j[0].shape)) for _, j in arr_h)
        res["half_load"] = {"req_per_s": round(0.5 * cap_rps, 2), "seconds": args.duration / 2, "requests": len(arr_h),
                            "Mvoxel_s": round(vox_h / 1e6 / wall_h, 1),
                            "latency_ms": {"p50": round(float(np.percentile(lat_h, 50)), 2),
                                           "p90": round(float(np.percentile(lat_h, 90)), 2),
                                           "p99": round(float(np.percentile(lat_h, 99)), 2)}}
    srv.stop()
    res["capacity"] = {"req_per_s": round(cap_rps, 2), "Mvoxel_s": round(cap_vox / 1e6 / cap_s, 1),
                       "how": f"144 requests submitted at once, at most {args.max_active} in flight"}
    if not args.no_plain:
        # the same trace on the plain path: one unfused call per request, FIFO, one worker
        lat_p = []
        done = threading.Event()
        q = [a for a in arr if a[0] <= args.plain_seconds]  # the trace's first plain_seconds

        def worker():
            t00 = time.perf_counter()
            for at, (v, meth, it, ch) in q:
                d = at - (time.perf_counter() - t00)
                if d > 0:
                    time.sleep(d)
                # PSFs centre-cropped to small volumes as the server does (RLdecon.m:264-270)
                pf = pk.crop_psf_to_data(psfs[ch], v.shape)
                if meth == "omw":
                    pb = pk.crop_psf_to_data(bps[ch], v.shape)
                    pk.decon_lucy_omw_function(v, pf, pb, it, Background=100, pad=True, ctx=ctx, flags=PLAIN)
                else:
                    pk.decon_lucy_function(v, pf, it, Background=100, pad=True, ctx=ctx, flags=PLAIN)
                lat_p.append((time.perf_counter() - t00 - at) * 1e3)
            done.set()

        progress(f"served: plain path on the first {args.plain_seconds:.0f} s of the trace ({len(q)} requests)")
        tp = time.perf_counter()
        th = threading.Thread(target=worker)
        th.start()
        th.join()
        pw = time.perf_counter() - tp
        busy_vox = sum(float(np.prod(j[0].shape)) for _, j in q)
        res["plain_gpu"] = {"Mvoxel_s": round(busy_vox / 1e6 / pw, 1), "seconds_of_trace": args.plain_seconds,
                            "requests": len(q),
                            "latency_ms": {"p50": round(float(np.percentile(lat_p, 50)), 2),
                                           "p99": round(float(np.percentile(lat_p, 99)), 2)},
                            "path": "one MATLAB-structured unfused call per request, OTFs rebuilt per call, FIFO"}
        res["capacity_speedup_vs_plain"] = round(res["capacity"]["Mvoxel_s"] / res["plain_gpu"]["Mvoxel_s"], 2)
    return res


def wl_served(args, torch, pk, L):
    out = served_section(args, torch, pk, half_load=True)
    res = base_result("served decon Mvoxel/s + p99 latency (OMW-2 / Biggs-15 mix)", out["Mvoxel_s"], "Mvoxel/s", args,
                      out["latency_ms"]["p50"], {"workload": out.pop("workload"), "requests": out["requests"]})
    res["served"] = out
    return res


def wl_large(args, torch, pk, L):
    from petakit5d_amd import synth
    shape = tuple(args.shape or (2048, 2048, 512))
    t_gen = time.perf_counter()
    It, psf = synth.large_blob_case(shape=shape)
    psf_b, _ = back_projector(psf, args.bp)
    t_gen = time.perf_counter() - t_gen
    ctx = pk.Context(0)
    nvox = float(np.prod(shape))
    out = {}
    tune = None if args.no_autotune else ctx.autotune(shape)
    for method, iters in (("omw", 2), ("rl", 10)):
        call = Caller(ctx, L, method, It, psf, psf_b, iters)
        ms, _, prof = timed(call, max(1, args.steps // 5), 1, torch, ctx, events_in_timed=args.events_in_timed)
        rl, passes, dev_ms = roofline(prof, max(1, args.steps // 5))
        out[f"{method}{iters}"] = {"Mvoxel_s": round(nvox / 1e6 / (ms / 1e3), 1), "ms_per_call": round(ms, 2),
                                   "roofline": rl, "passes": passes,
                                   "device_GB": round(ctx.device_bytes() / 2 ** 30, 1)}
        del call
    res = base_result("deconvolved Mvoxel/s (OMW, 2 it, 2048x2048x512)", out["omw2"]["Mvoxel_s"], "Mvoxel/s", args,
                      out["omw2"]["ms_per_call"], {
                          "workload": f"{shape[0]}x{shape[1]}x{shape[2]} uint16 (2^31 voxels), 500k blobs, 64x64x101 "
                                      "PSF, fully resident in HBM; OMW-2 and Biggs-10", "volume": list(shape)})
    res["large"] = out
    if tune:
        res["autotune"] = tune
    res["gen_s"] = round(t_gen, 1)
    if not args.no_check:
        po = oracle()
        sub = np.asfortranarray(It[:256, :256, :256].cpu().numpy())
        Jg, _ = pk.decon_lucy_omw_function(sub, psf, psf_b, 2, Background=100, save16bit=False, ctx=ctx)
        J64, _ = po.decon_lucy_omw_function(sub, psf, psf_b, 2, Background=100, save16bit=False, precision="f64")
        J32, _ = po.decon_lucy_omw_function(sub, psf, psf_b, 2, Background=100, save16bit=False)
        res["parity_256cube"] = {"rel_l2_vs_fp64": rel_l2(Jg, J64), "rel_l2_vs_cpu_fp32": rel_l2(Jg, J32),
                                 "cpu_fp32_vs_fp64": rel_l2(J32, J64)}
    return res


def wl_copy(args, torch, pk, L):
    """A plain device copy of known size: the calibration kernel for the HBM
    byte counters (FETCH_SIZE / WRITE_SIZE) and the achievable copy rate."""
    nbytes = 1 << 30
    gbps = copy_bandwidth(torch, nbytes, max(1, args.steps))
    res = base_result("device copy GB/s", gbps, "GB/s", args, 2.0 * nbytes / gbps / 1e6, {
        "workload": "torch copy_ of a 1 GiB float32 buffer (1 GiB read + 1 GiB written per step)", "bytes": nbytes})
    res["dtype"] = "f32"
    return res


def wl_probes(args, torch, pk, L):
    """Copy probes (pk_probe_copy): the pitched half spectrum of a transform
    copied with each pass's access pattern and no arithmetic -- the HBM
    ceilings the passes are compared with.  Sequential 16-byte copy, the
    dim-2 column pattern (y passes) and the dim-3 column pattern (z_conv) at
    several tile widths T and values per thread E, XCD-aware block order (and
    the linear order beside it), and the dim-3 pattern with a second input
    read (z_conv's three streams), at 512x512x256 and 2048x2048x512; each copy
    timed back to back and cold (after a 1 GiB read that flushes the
    memory-side cache).  Column copies store as the strided passes do
    (nontemporally for tiles of 16 or more columns).  `value` is the cold dim-3 pattern with a second input at the
    shape z_conv runs with at 512x512x256 (T = 16, E = 16)."""
    ctx = pk.Context(0)
    reps = max(5, args.steps)
    out = {}
    cases = {(512, 512, 256): [(0, 0, 0), (1, 8, 16), (1, 16, 16), (1, 32, 16), (1, 16, 32), (1, 64, 32),
                               (2, 8, 16), (2, 16, 16), (2, 32, 16), (2, 64, 16), (2, 16, 8), (2, 32, 8),
                               (3, 16, 16), (3, 32, 16)],
             (2048, 2048, 512): [(0, 0, 0), (1, 8, 16), (1, 8, 32), (1, 16, 32), (2, 8, 16), (2, 16, 16), (2, 32, 16),
                                 (2, 16, 32), (2, 32, 32), (2, 64, 32), (3, 16, 8), (3, 16, 16), (3, 32, 16),
                                 (3, 32, 32)]}
    names = {0: "sequential", 1: "dim2_columns", 2: "dim3_columns", 3: "dim3_columns_two_inputs"}
    for dims, cs in cases.items():
        key = "x".join(map(str, dims))
        rec = {}
        for pat, T, E in cs:
            progress(f"probe {key} {names[pat]} T={T} E={E}")
            ms, nb = ctx.probe_copy(pat, dims, T or 16, E or 16, True, reps)
            key_c = names[pat] if pat == 0 else f"{names[pat]}_T{T}_E{E}"
            rec[key_c] = {"ms": round(ms, 4), "GB_s": round(nb / ms / 1e6, 1)}
            ms2, _ = ctx.probe_copy(pat, dims, T or 16, E or 16, False, reps) if pat else (None, None)
            if ms2:
                rec[key_c]["GB_s_linear_order"] = round(nb / ms2 / 1e6, 1)
            ms3, _ = ctx.probe_copy(pat, dims, T or 16, E or 16, True, reps, cold=True)
            rec[key_c]["GB_s_cold"] = round(nb / ms3 / 1e6, 1)
        out[key] = rec
    ctx.close()
    v = out["512x512x256"]["dim3_columns_two_inputs_T16_E16"]["GB_s_cold"]
    res = base_result("strided-pattern copy GB/s (dim-3 columns + second input, T=16, E=16, 512x512x256, cold)", v,
                      "GB/s", args, out["512x512x256"]["dim3_columns_two_inputs_T16_E16"]["ms"], {
                          "workload": "copy of the pitched half spectrum with each pass's access pattern, no "
                                      "arithmetic (read + written bytes of the valid columns)"})
    res["probes"] = out
    res["dtype"] = "f32"
    return res


def wl_deskew(args, torch, pk, L):
    """deskewFrame3D of a decon result (single, 512x512x256, the default
    workload's size) at the reference defaults SkewAngle 32.45, dz 0.3,
    xyPixelSize 0.108: device in, device out, one pk_deskew_frame3d per step;
    kernel time from HIP events on the engine's stream; CPU baseline the
    oracle's numpy resampling on a 64-plane slab of the same volume."""
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    shape, geo = (512, 512, 256), (32.45, 0.3, 0.108)
    g = torch.Generator(device="cuda").manual_seed(5)
    vol = torch.rand((shape[2], shape[1], shape[0]), device="cuda", generator=g).permute(2, 1, 0)
    od = pk.deskew.deskew_out_size(shape, *geo)
    out = torch.empty((od[2], od[1], od[0]), device="cuda").permute(2, 1, 0)
    ctx = pk.default_context()
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)

    def call():
        ctx.check(L.pk_deskew_frame3d(ctx.handle, ctypes.c_void_p(vol.data_ptr()), _lib_i64x3(shape),
                                      ctypes.c_double(geo[0]), ctypes.c_double(geo[1]), ctypes.c_double(geo[2]),
                                      0, 0, ctypes.c_void_p(out.data_ptr())))

    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for a, b in evs:
        a.record(stream)
        call()
        b.record(stream)
    torch.cuda.synchronize()
    kms = sum(a.elapsed_time(b) for a, b in evs) / len(evs)
    ctx.set_stream(None)
    nin, nout = vol.numel(), out.numel()
    nbytes = 4.0 * (nin + nout)
    res = base_result("deskewed Mvoxel/s (deskewFrame3D, linear)", nout / ms / 1e3, "Mvoxel/s", args, ms, {
        "workload": "deskewFrame3D of a 512x512x256 single volume, SkewAngle 32.45, dz 0.3, xyPixelSize 0.108, "
                    "linear, no crop; device-resident in and out",
        "in": list(shape), "out": list(od)})
    res["roofline"] = {"bound": "hbm", "kernel": "k_deskew", "achieved": round(nbytes / kms / 1e6, 1),
                       "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(nbytes / kms / 1e6 / HBM_PEAK_GBPS, 4),
                       "traffic": None, "algorithmic_bytes": nbytes, "kernel_ms": round(kms, 4)}
    if not args.no_cpu_baseline:
        import deskew_oracle
        slab = vol[:, :, :64].permute(2, 1, 0).cpu().numpy().transpose(2, 1, 0)
        t0 = time.perf_counter()
        ref = deskew_oracle.deskew(slab, *geo)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(ref.size / dt / 1e6, 2), "unit": "Mvoxel/s", "cores": 1, "kind": "port",
                               "sample": f"512x512x64 slab, numpy float64, {dt:.2f} s"}
    return res


def wl_rotate(args, torch, pk, L):
    """rotateFrame3D of a deskewed 512x512x256 stage-scan volume (the deskew
    workload's output, 512x1110x256) at SkewAngle 32.45, zAniso = sind(32.45)
    0.3 / 0.108, automatic crop: one pk_affine_warp3d per step on device
    buffers, kernel time from HIP events on the engine's stream."""
    import ctypes
    from petakit5d_amd.rotate import rotate_transform
    shape, geo = (512, 512, 256), (32.45, 0.3, 0.108)
    g = torch.Generator(device="cuda").manual_seed(6)
    vol = torch.rand((shape[2], shape[1], shape[0]), device="cuda", generator=g).permute(2, 1, 0)
    ds = pk.deskewFrame3D(vol, *geo)
    dsh = tuple(int(n) for n in ds.shape)
    zan = math.sin(geo[0] * math.pi / 180) * geo[1] / geo[2]
    T, od = rotate_transform(dsh, geo[0], zan, proj=ds.amax(dim=0).cpu().numpy())
    inv = np.ascontiguousarray(np.linalg.inv(T))
    out = torch.empty((od[2], od[1], od[0]), device="cuda").permute(2, 1, 0)
    ctx = pk.default_context()
    stream = torch.cuda.current_stream()
    ctx.set_stream(stream.cuda_stream)

    def call():
        ctx.check(L.pk_affine_warp3d(ctx.handle, ctypes.c_void_p(ds.data_ptr()), _lib_i64x3(dsh),
                                     inv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _lib_i64x3(od),
                                     ctypes.c_void_p(out.data_ptr())))

    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for a, b in evs:
        a.record(stream)
        call()
        b.record(stream)
    torch.cuda.synchronize()
    kms = sum(a.elapsed_time(b) for a, b in evs) / len(evs)
    ctx.set_stream(None)
    nbytes = 4.0 * (ds.numel() + out.numel())
    res = base_result("rotated Mvoxel/s (rotateFrame3D, linear)", out.numel() / ms / 1e3, "Mvoxel/s", args, ms, {
        "workload": "rotateFrame3D of a deskewed 512x1110x256 single volume, SkewAngle 32.45, zAniso 1.49, "
                    "automatic crop; device-resident in and out", "in": list(dsh), "out": list(od)})
    res["roofline"] = {"bound": "hbm", "kernel": "k_affine3d", "achieved": round(nbytes / kms / 1e6, 1),
                       "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(nbytes / kms / 1e6 / HBM_PEAK_GBPS, 4),
                       "traffic": None, "algorithmic_bytes": nbytes, "kernel_ms": round(kms, 4)}
    return res


def _lib_i64x3(v):
    from petakit5d_amd._lib import i64x3
    return i64x3(v)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported. Volumes and chunks are "
              "independent units: a multi-GPU deployment runs one process and context per GPU, each "
              "with its own server, and assigns whole requests or chunks to GPUs", file=sys.stderr)
        sys.exit(2)
    import torch

    import petakit5d_amd as pk
    from petakit5d_amd import _lib

    L = _lib.load()
    res = {"omw": wl_omw, "rl128": wl_rl128, "batched": wl_batched, "served": wl_served, "large": wl_large,
           "copy": wl_copy, "deskew": wl_deskew, "rotate": wl_rotate, "probes": wl_probes}[
        args.workload](args, torch, pk, L)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

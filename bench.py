#!/usr/bin/env python3
"""One-GPU benchmark of the OMW deconvolution hot path.

Workload: decon_lucy_omw_function with 2 iterations on a 512x512x256 uint16
volume (20k Gaussian blobs, smooth background, blurred by a 64x64x101 PSF
sheared 31.5 deg, Poisson noise), Background = 100, OTFs cached on the device,
input and uint16 output resident in HBM.  One step = one decon call.

Prints ONE JSON line.  `value` is deconvolved Mvoxel/s on one GPU; `roofline`
is for the pass with the largest share of device time (algorithmic bytes per
launch / average launch time, HIP events on the engine's stream);
`cpu_baseline` times the oracle (the CPU single-precision restatement) on the
host cores for the same call.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--shape", type=int, nargs=3, default=[512, 512, 256])
    ap.add_argument("--iters", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the parity check against the oracle")
    return ap.parse_args()


def col_major_cuda(a, torch):
    return torch.from_numpy(np.ascontiguousarray(a.transpose(2, 1, 0))).cuda().permute(2, 1, 0)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(volumes are independent units, see the sharding notes)", file=sys.stderr)
        sys.exit(2)
    import torch

    import petakit5d_amd as pk
    from petakit5d_amd import _lib, synth

    shape = tuple(args.shape)
    t_gen = time.perf_counter()
    I, psf = synth.config2(shape=shape)
    psf_b, bp_kind = back_projector(psf)
    t_gen = time.perf_counter() - t_gen

    ctx = pk.Context(0)
    L = _lib.load()
    It = col_major_cuda(I, torch)
    out = torch.empty((shape[2], shape[1], shape[0]), dtype=torch.uint16, device="cuda").permute(2, 1, 0)
    vin = _lib.PkVol(ctypes.c_void_p(It.data_ptr()), _lib.PK_UINT16, _lib.PK_MEM_DEVICE, (ctypes.c_int64 * 3)(*shape))
    vout = _lib.PkVol(ctypes.c_void_p(out.data_ptr()), _lib.PK_UINT16, _lib.PK_MEM_DEVICE,
                      (ctypes.c_int64 * 3)(*shape))
    opts = _lib.PkOpts()
    L.pk_opts_default(ctypes.byref(opts))
    opts.background = 100.0
    pf = np.asfortranarray(psf, dtype=np.float32)
    pb = np.asfortranarray(psf_b, dtype=np.float32)
    fd, bd = _lib.i64x3(pf.shape), _lib.i64x3(pb.shape)
    torch.cuda.synchronize()

    def step():
        ctx.check(L.pk_decon_lucy_omw(ctx.handle, ctypes.byref(vin), ctypes.c_void_p(pf.ctypes.data), fd,
                                      ctypes.c_void_p(pb.ctypes.data), bd, args.iters, ctypes.byref(opts),
                                      ctypes.byref(vout), None))

    for _ in range(max(args.warmup, 1)):  # the first call builds and caches the two OTFs
        step()
    torch.cuda.synchronize()
    ctx.profile(True)
    lat = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        step()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = ctx.profile_read()
    ctx.profile(False)
    ms = (t1 - t0) / args.steps * 1e3
    nvox = float(np.prod(shape))
    value = nvox / 1e6 / (ms / 1e3)

    # dominant pass: largest total device time
    name, (launches, tot_ms, bpl) = max(prof.items(), key=lambda kv: kv[1][1])
    avg_ms = tot_ms / launches
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    dev_ms = sum(v[1] for v in prof.values()) / args.steps
    N = nvox
    S = (shape[0] // 2 + 1) * shape[1] * shape[2]
    model_bytes = args.iters * (24 * N + 176 * S) + 12 * N  # per call, OTFs cached
    passes = {k: {"launches_per_step": v[0] / args.steps, "avg_ms": round(v[1] / v[0], 4),
                  "GBps": round(v[2] / (v[1] / v[0] * 1e-3) / 1e9, 1)} for k, v in sorted(prof.items())}

    result = {
        "metric": "deconvolved Mvoxel/s (OMW, 2 it)",
        "value": round(value, 2),
        "unit": "Mvoxel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"OMW decon, {args.iters} iterations, {shape[0]}x{shape[1]}x{shape[2]} uint16, "
                        "20k blobs + sinusoidal background, 64x64x101 PSF sheared 31.5 deg, Poisson; "
                        "OTFs cached, input/output resident in HBM; one step = one decon call",
            "volume": list(shape),
            "psf": [64, 64, 101],
            "iterations": args.iters,
            "back_projector": bp_kind,
            "background": 100,
            "output": "uint16",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": name,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": None,
        },
        "call_model": {
            "bytes_per_call": model_bytes,
            "GBps_vs_wall": round(model_bytes / (ms * 1e-3) / 1e9, 1),
            "frac_vs_wall": round(model_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
            "device_ms_per_step": round(dev_ms, 4),
        },
        "passes": passes,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "gen_s": round(t_gen, 1),
    }

    if not args.no_check or not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pk_oracle

        pk_oracle.build()
    if not args.no_check:
        Jg, _ = pk.decon_lucy_omw_function(I, psf, psf_b, args.iters, Background=100, save16bit=False, ctx=ctx)
    if not args.no_cpu_baseline or not args.no_check:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        tc = time.perf_counter()
        Jc, _ = pk_oracle.decon_lucy_omw_function(I, psf, psf_b, args.iters, Background=100, save16bit=False)
        tc = time.perf_counter() - tc
        if not args.no_cpu_baseline:
            result["cpu_baseline"] = {"value": round(nvox / 1e6 / tc, 3), "unit": "Mvoxel/s", "cores": threads,
                                      "kind": "port",
                                      "sample": f"one full {shape[0]}x{shape[1]}x{shape[2]} OMW-{args.iters} call "
                                                "(oracle FP32, full-spectrum FFTs, OTF build included), "
                                                f"{tc:.1f} s"}
        if not args.no_check:
            a, b = Jg.astype(np.float64), Jc.astype(np.float64)
            result["parity"] = {"rel_l2_vs_cpu_fp32": float(np.linalg.norm(a - b) / np.linalg.norm(b))}
    print(json.dumps(result), flush=True)


def back_projector(psf):
    """OMW back projector of the PSF (omw_backprojector_generation, alpha 0.005)."""
    try:
        from petakit5d_amd.backprojector import omw_backprojector_generation
    except ImportError:
        return np.ascontiguousarray(psf[::-1, ::-1, ::-1]), "flip(PSF) (traditional RL back projector)"
    b, _, _, _ = omw_backprojector_generation(psf, 0.005, None, OTFCumThresh=0.9, hannWinBounds=[0.8, 1])
    return b, "OMW, alpha 0.005, OTFCumThresh 0.9, hannWinBounds [0.8 1]"


if __name__ == "__main__":
    main()

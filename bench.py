"""Benchmark of the 2D NSST decompose -> hard threshold -> reconstruct path.

Workload (one step): the plain single-image NSST hard-threshold denoise of one
512 x 512 FP32 image resident in HBM -- maxflat a trous pyramid with 4 levels,
2^[3 3 4 4] = [8 8 16 16] shearing directions, Meyer filter sizes
[32 32 16 16], thresholds [0 3 4]*sigma*s per band (sigma = 20) -- through the
C ABI (shx_denoise: shear_trans, then the inverse with the thresholds fused
into its first read of each band).  Input: 12 random filled ellipses plus
N(0, 20^2) noise, seeded.

Prints one JSON line.  ``value`` is Mpixel/s of whole steps.  ``roofline`` is
for the dominant kernel (longest mean launch duration) measured with HIP
events on the launch stream; ``path_roofline`` applies the fixed band-stack
byte model of the whole step (400 B per pixel: input, 49 bands written, 49
read, output).  ``cpu_baseline`` times the FP64 numpy oracle on the same
workload for a bounded number of images on one host core.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "NSST dec+thresh+rec Mpixel/s @512² FP32, batched vs plain; adjoint ms/iter"


def cpu_baseline(L, sigma, budget_s):
    import numpy as np

    from oracle import nsst as o
    from shearcuda_amd.synth import noisy_ellipses, white_noise
    H = o.make_dictionary(L)
    sc = o.nsst_scalars(H, white_noise(L))
    _, x = noisy_ellipses(L, sigma=sigma, seed=0)
    n = 0
    t0 = time.perf_counter()
    while True:
        d = o.nsst_dec2(x, H)
        o.nsst_rec2(o.nsst_HT(d, sigma, (0, 3, 4), sc), H)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    return {"value": round(L * L * n / el / 1e6, 4), "unit": "Mpixel/s", "cores": 1, "kind": "port",
            "sample": "%d image(s) of %dx%d through the FP64 numpy oracle (dec + nsst_HT + rec), %.1f s" % (
                n, L, L, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU oracle work (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the NSST path has no multi-GPU form here); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from shearcuda_amd import api
    from shearcuda_amd._lib import check, lib, ptr_array
    from shearcuda_amd.synth import noisy_ellipses, white_noise

    L, sigma = args.size, 20.0
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    shear = api.make_dictionary(L)
    _, x = noisy_ellipses(L, sigma=sigma, seed=0)
    xg = torch.from_numpy(x.astype(np.float32)).to(dev)
    sc = api.nsst_scalars_cuda(shear, torch.from_numpy(white_noise(L).astype(np.float32)).to(dev))
    Ct = shear.alloc_coeffs()
    out = torch.empty_like(xg)

    for _ in range(args.warmup):
        api.denoise_cuda(xg, shear, sigma, sc, Ct=Ct, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        api.denoise_cuda(xg, shear, sigma, sc, Ct=Ct, out=out)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms = el * 1e3 / args.steps
    value = L * L * args.steps / el / 1e6

    # per-kernel durations with HIP events on the launch stream
    MAXS, NL = 64, 32
    stage_ms = (ctypes.c_double * MAXS)()
    stage_b = (ctypes.c_double * MAXS)()
    names = ctypes.create_string_buffer(MAXS * NL)
    nst = ctypes.c_int()
    scv = np.concatenate([[float(sc[0])]] + [np.asarray(s, dtype=np.float64) for s in sc[1:]])
    T = np.array([0.0, 3.0, 4.0])
    reps = max(10, args.steps)
    check(lib.shx_denoise_profile(shear.handle, ctypes.c_void_p(xg.data_ptr()), ptr_array([c.data_ptr() for c in Ct]),
                                  sigma, T.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  scv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.c_void_p(out.data_ptr()),
                                  reps, MAXS, stage_ms, stage_b, names, NL, ctypes.byref(nst),
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    stages = []
    for i in range(nst.value):
        nm = names.raw[i * NL:(i + 1) * NL].split(b"\0", 1)[0].decode()
        stages.append({"name": nm, "ms": stage_ms[i], "bytes": stage_b[i],
                       "GBs": stage_b[i] / (stage_ms[i] * 1e-3) / 1e9 if stage_ms[i] > 0 else 0.0})
    dom = max(stages, key=lambda s: s["ms"])
    roof = {"kernel": dom["name"], "bound": "hbm", "achieved": round(dom["GBs"], 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(dom["GBs"] / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_launch": dom["bytes"], "ms_per_launch": round(dom["ms"], 5)}
    path_bytes = 400.0 * L * L
    path_gbs = path_bytes / (ms * 1e-3) / 1e9
    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpixel/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain 2D NSST hard-threshold denoise (dec+thresh+rec), one %dx%d FP32 image per step, "
                               "maxflat 4 levels, directions [8,8,16,16], Meyer sizes [32,32,16,16], T=[0,3,4], "
                               "sigma=20" % (L, L), "L": L, "images_per_step": 1, "form": "plain"},
        "roofline": roof,
        "path_roofline": {"bytes_per_step": path_bytes, "achieved": round(path_gbs, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(path_gbs / HBM_PEAK_GBS, 4)},
        "kernels": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in s.items()} for s in stages],
    }
    if args.cpu_budget > 0:
        res["cpu_baseline"] = cpu_baseline(L, sigma, args.cpu_budget)
    shear.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

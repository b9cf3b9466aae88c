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
read, output).  ``cpu_baseline`` times the FP64 numpy oracle (the CPU
restatement of shearlet_toolbox_1) on the same workload for a bounded time,
on all host cores (one process per core) with the one-core figure beside it.
"""
import argparse
import csv
import ctypes
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def path_flops_per_px(L):
    """Fixed FLOP model of the 2D denoise per pixel (maxflat level pyramid
    8480, 104 real FFTs at 2.5 log2(L^2) each, spectral products 289,
    threshold and accumulate 100): the VALU counterpart of the 400 B/px byte
    model.  Implementation-independent, so the DCT-domain pyramid is priced
    as the level pyramid it replaces."""
    import math
    return 8480 + 104 * 2.5 * math.log2(L * L) + 289 + 100


def valu_frac(L, npx, seconds):
    return round(path_flops_per_px(L) * npx / seconds / 1e12 / FP32_PEAK_TFLOPS, 4)
METRIC = "NSST dec+thresh+rec Mpixel/s @512² FP32, batched vs plain; adjoint ms/iter"


def _oracle_setup(L, sigma):
    from oracle import nsst as o
    from shearcuda_amd.synth import noisy_ellipses, white_noise
    H = o.make_dictionary(L)
    sc = o.nsst_scalars(H, white_noise(L))
    _, x = noisy_ellipses(L, sigma=sigma, seed=0)
    return H, sc, x


def _oracle_worker(L, sigma, seconds, ready, out):
    """Worker of the all-core baseline: builds its dictionary, waits for the
    others, then runs FP64 oracle denoise steps (dec + nsst_HT + rec) for
    `seconds`; puts (steps, elapsed) on out."""
    from oracle import nsst as o
    H, sc, x = _oracle_setup(L, sigma)
    ready.wait()
    n = 0
    t0 = time.perf_counter()
    while True:
        d = o.nsst_dec2(x, H)
        o.nsst_rec2(o.nsst_HT(d, sigma, (0, 3, 4), sc), H)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            out.put((n, el))
            return


def cpu_baseline(L, sigma, budget_s, workers=None):
    """CPU restatement of shearlet_toolbox_1 (the FP64 numpy oracle: nsst_dec2,
    nsst_HT, nsst_rec2) on the same workload, timed on one core and on all
    host cores (one process per core, each denoising its own copy of the
    image, all started together after their set-up).  Runs before the process
    touches the GPU, so the workers are forks of a CPU-only process."""
    import multiprocessing as mp

    from oracle import nsst as o
    H, sc, x = _oracle_setup(L, sigma)
    n = 0
    t0 = time.perf_counter()
    while True:
        d = o.nsst_dec2(x, H)
        o.nsst_rec2(o.nsst_HT(d, sigma, (0, 3, 4), sc), H)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s / 2:
            break
    single = {"value": round(L * L * n / el / 1e6, 4), "unit": "Mpixel/s", "cores": 1,
              "sample": "%d image(s) of %dx%d on one core, %.1f s" % (n, L, L, el)}
    del H
    # all cores: the box's CPU share is at most 16 logical CPUs per GPU
    workers = workers or max(1, min(16, len(os.sched_getaffinity(0))))
    ctx = mp.get_context("fork")
    ready = ctx.Barrier(workers)
    out = ctx.Queue()
    procs = [ctx.Process(target=_oracle_worker, args=(L, sigma, budget_s / 2, ready, out)) for _ in range(workers)]
    for p in procs:
        p.start()
    res = [out.get() for _ in procs]
    for p in procs:
        p.join()
    tot = sum(r[0] for r in res)
    rate = sum(r[0] / r[1] for r in res)     # the workers run concurrently
    el = max(r[1] for r in res)
    return {"value": round(L * L * rate / 1e6, 4), "unit": "Mpixel/s", "cores": workers, "kind": "port",
            "label": "CPU restatement of shearlet_toolbox_1 (FP64 numpy oracle: nsst_dec2 + nsst_HT + nsst_rec2)",
            "sample": "%d images of %dx%d over %d processes (one per core) in %.1f s, dictionary set-up excluded"
                      % (tot, L, L, workers, el),
            "single_core": single, "host": host_cpu()}


def host_cpu():
    """The host the CPU baseline ran on: logical CPUs and the model name."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu": model}


def _timed(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def plain_setup(L, sigma, seed=0):
    import numpy as np
    import torch

    from shearcuda_amd import api
    from shearcuda_amd.synth import noisy_ellipses, white_noise
    shear = api.make_dictionary(L)
    _, x = noisy_ellipses(L, sigma=sigma, seed=seed)
    xg = torch.from_numpy(x.astype(np.float32)).cuda()
    sc = api.nsst_scalars_cuda(shear, torch.from_numpy(white_noise(L).astype(np.float32)).cuda())
    return shear, xg, sc


def dictionary_build_times(L):
    """Seconds to build the default dictionary (spatial filters, spectra,
    pyramid masks, workspaces) with the spectra computed on the device and on
    the host."""
    import torch

    from shearcuda_amd import api
    for where in ("device", "host"):  # warm-up: module loading and first allocations
        api.make_dictionary(64, spectra=where).close()
    out = {}
    for where in ("device", "host"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        shear = api.make_dictionary(L, spectra=where)
        torch.cuda.synchronize()
        out[where] = round(time.perf_counter() - t0, 3)
        shear.close()
    return out


def profile_plain(shear, xg, sc, sigma, reps, coeffs=False):
    """Per-kernel mean durations of one plain denoise step (HIP events)."""
    import numpy as np
    import torch

    from shearcuda_amd._lib import check, lib, ptr_array
    cells = ptr_array([c.data_ptr() for c in shear.alloc_coeffs()]) if coeffs else None
    out = torch.empty_like(xg)
    MAXS, NL = 64, 32
    stage_ms = (ctypes.c_double * MAXS)()
    stage_b = (ctypes.c_double * MAXS)()
    names = ctypes.create_string_buffer(MAXS * NL)
    nst = ctypes.c_int()
    scv = np.concatenate([[float(sc[0])]] + [np.asarray(s, dtype=np.float64) for s in sc[1:]])
    T = np.array([0.0, 3.0, 4.0])
    check(lib.shx_denoise_profile(shear.handle, ctypes.c_void_p(xg.data_ptr()), cells, sigma, T.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  scv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.c_void_p(out.data_ptr()),
                                  reps, MAXS, stage_ms, stage_b, names, NL, ctypes.byref(nst),
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    stages = []
    for i in range(nst.value):
        nm = names.raw[i * NL:(i + 1) * NL].split(b"\0", 1)[0].decode()
        stages.append({"name": nm, "ms": stage_ms[i], "bytes": stage_b[i],
                       "GBs": stage_b[i] / (stage_ms[i] * 1e-3) / 1e9 if stage_ms[i] > 0 else 0.0})
    return stages


def apply_schedule(shear, sched):
    from shearcuda_amd import api
    api.set_plan_options(shear, dct_front=sched["dct_front"], dct_back=sched["dct_back"])
    api.set_schedule(shear, sched["pyramid_variant"], sched["group"], col_tile_dec=sched["col_tile_dec"],
                     col_tile_rec=sched["col_tile_rec"])


def bench_plain(L, sigma, steps, warmup, sched=None):
    """ms per step of the plain denoise graph returning the image only (the
    timed step), ms per step of the same step also writing the unthresholded
    coefficients, the per-kernel profile of the timed step and the schedule
    (autotuned unless given)."""
    from shearcuda_amd import api
    shear, xg, sc = plain_setup(L, sigma)
    if sched is None:
        sched = api.autotune(shear)
    else:
        apply_schedule(shear, sched)
    g, _, out = api.denoise_graph(xg, shear, sigma, sc, coeffs=False)
    el = _timed(g.launch, steps, warmup)
    gc, Ct, outc = api.denoise_graph(xg, shear, sigma, sc)
    elc = _timed(gc.launch, steps, warmup)
    stages = profile_plain(shear, xg, sc, sigma, max(20, steps))
    g.close()
    gc.close()
    shear.close()
    return el * 1e3 / steps, elc * 1e3 / steps, stages, sched


# mixed-size batch: 48 x 256^2, 32 x 512^2, 16 x 1024^2, sigma ~ U[5, 40]
BATCH_MIX = [(256, 48), (512, 32), (1024, 16)]


def bench_batched(steps, warmup, mix=BATCH_MIX):
    """One graph launch denoises every image of the mix (per-image sigma)."""
    import numpy as np
    import torch

    from shearcuda_amd import api
    from shearcuda_amd.synth import noisy_ellipses, white_noise
    rng = np.random.default_rng(2024)
    plans, npx, seed = [], 0, 100
    for L, n in mix:
        shear = api.make_dictionary(L)
        sc = api.nsst_scalars_cuda(shear, torch.from_numpy(white_noise(L).astype(np.float32)).cuda())
        sig = rng.uniform(5.0, 40.0, size=n)
        xs = np.stack([noisy_ellipses(L, sigma=float(sig[i]), seed=seed + i)[1] for i in range(n)]).astype(np.float32)
        seed += n
        b = api.Batch(shear, n)
        sched = api.autotune(shear, b)
        plans.append(dict(sched=sched, L=L, n=n, shear=shear, sc=sc, sig=sig, b=b, x=torch.from_numpy(xs).cuda(),
                          th=api.batch_thresholds(shear, sig, sc), out=torch.empty((n, L, L), device="cuda")))
        npx += n * L * L
    # timed: the denoised images only; beside it the same launch also writing
    # every image's unthresholded coefficients
    g = api.capture(lambda: [p["b"].denoise(p["x"], None, p["th"], p["out"]) for p in plans])
    el = _timed(g.launch, steps, warmup)
    batched = npx * steps / el / 1e6
    g.close()
    for p in plans:
        p["c"] = p["b"].alloc_coeffs()
    gcf = api.capture(lambda: [p["b"].denoise(p["x"], p["c"], p["th"], p["out"]) for p in plans])
    elc = _timed(gcf.launch, steps, warmup)
    gcf.close()
    for p in plans:
        del p["c"]
    # the same images one at a time through the plain form (one graph each)
    singles = []
    for p in plans:
        for i in range(p["n"]):
            gi, _, _ = api.denoise_graph(p["x"][i], p["shear"], float(p["sig"][i]), p["sc"], coeffs=False)
            singles.append(gi)
    el1 = _timed(lambda: [gi.launch() for gi in singles], max(2, steps // 2), 1)
    plain = npx * max(2, steps // 2) / el1 / 1e6
    for gi in singles:
        gi.close()
    for p in plans:
        p["b"].close()
        p["shear"].close()
    return {"images": sum(n for _, n in mix), "pixels": npx, "mix": [[L, n] for L, n in mix],
            "ms_per_launch": round(el * 1e3 / steps, 4), "Mpixel_s": round(batched, 2),
            "Mpixel_s_with_coeffs": round(npx * steps / elc / 1e6, 2),
            "plain_image_by_image_Mpixel_s": round(plain, 2), "speedup_vs_plain": round(batched / plain, 3),
            "path_roofline_frac": round(400.0 * npx / (el / steps) / 1e9 / HBM_PEAK_GBS, 4),
            "path_valu_frac": round(sum(path_flops_per_px(L) * L * L * n for L, n in mix) / (el / steps) / 1e12
                                    / FP32_PEAK_TFLOPS, 4),
            "schedules": {str(p["L"]): p["sched"] for p in plans}}


def bench_separate(L=512, reps=10):
    """The shipped separation loop (separate.m) with the parameters of
    separation_demo01_512_gpu.m:37,78: filters 80x80, directions [8 8 16 16],
    4 wavelet levels, 10 iterations, stop 3, gamma 3, weights [.1 .1 1.5 1.5];
    seconds per call, setup included, on an autotuned dictionary.  The loop is
    host-driven (about half of a call is launch and host work), so the median
    of `reps` individually timed calls after two warm-up calls is reported."""
    import statistics

    import numpy as np
    import torch

    from shearcuda_amd import api, separation
    from shearcuda_amd.synth import separation_input
    shear = api.make_dictionary(L, dcomp=(3, 3, 4, 4), dsize=(80, 80, 80, 80))
    api.autotune(shear)
    img = torch.from_numpy(separation_input(L, seed=1, sigma=10.0)[2].astype(np.float32)).cuda()
    run = lambda: separation.separate(img, 4, 10, 3.0, 3.0, [0.1, 0.1, 1.5, 1.5], shear)
    for _ in range(2):
        run()
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    native = lambda: separation.separate_native(img, 4, 10, 3.0, 3.0, [0.1, 0.1, 1.5, 1.5], shear)
    for _ in range(2):
        native()
    tn = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        native()
        torch.cuda.synchronize()
        tn.append(time.perf_counter() - t0)
    shear.close()
    s = statistics.median(times)
    sn = statistics.median(tn)
    return {"workload": "separate(img, 4, 10, 3, 3, [.1 .1 1.5 1.5], shear) on one 512x512 FP32 image, filters "
                        "80x80, directions [8,8,16,16]", "s_per_call": round(s, 4),
            "s_per_call_min": round(min(times), 4), "ms_per_iteration": round(s * 1e3 / 10, 3),
            "native_s_per_call": round(sn, 4), "native_ms_per_iteration": round(sn * 1e3 / 10, 3)}


def _adjoint_setup(iters=10, nimg=8, L=512):
    """The unrolled separation of bench_adjoint: (forward, forward + backward,
    objects to keep alive)."""
    import numpy as np
    import torch

    from shearcuda_amd import api, separation
    from shearcuda_amd.synth import separation_input, white_noise
    from shearcuda_amd.wavelets import MakeONFilter
    sigma, levels = 10.0, 3
    h = MakeONFilter("Daubechies", 4)
    shear = api.make_dictionary(L)
    plan = api.Batch(shear, nimg)
    api.autotune(shear, plan)
    noise = torch.from_numpy(white_noise(L).astype(np.float32)).cuda()
    s_sh = api.band_thresholds(shear, 1.0, api.nsst_scalars_cuda(shear, noise), (1.0, 1.0, 1.0))
    s_w = separation.wavelet_scalars(noise, h, levels)
    data = [separation_input(L, seed=300 + i, sigma=sigma) for i in range(nimg)]
    x = torch.from_numpy(np.stack([d[2] for d in data]).astype(np.float32)).cuda()
    gt = torch.from_numpy(np.stack([d[1] for d in data]).astype(np.float32)).cuda()
    th_sh = torch.tensor(3 * sigma * s_sh, dtype=torch.float32, device="cuda", requires_grad=True)
    th_w = torch.tensor(3 * sigma * s_w, dtype=torch.float32, device="cuda", requires_grad=True)

    def fwd():
        with torch.no_grad():
            separation.unrolled_separation(x, th_sh, th_w, plan, h, levels, iters)

    def fwd_bwd():
        xr = x.clone().requires_grad_(True)
        c, _ = separation.unrolled_separation(xr, th_sh, th_w, plan, h, levels, iters)
        torch.mean((c - gt) ** 2).backward()

    return fwd, fwd_bwd, (shear, plan)


def bench_adjoint(iters=10, nimg=8, L=512, reps=3):
    """Unrolled separation, 8 x 512^2 images: ms per unrolled iteration of the
    forward pass and of forward + backward (adjoint kernels)."""
    fwd, fwd_bwd, (shear, plan) = _adjoint_setup(iters, nimg, L)
    # two untimed calls each: the caching allocator's pool for the coefficient
    # stacks kept for the backward (one per iteration) fills over the first two
    t_f = _timed(fwd, reps, 2)
    t_fb = _timed(fwd_bwd, reps, 2)
    plan.close()
    shear.close()
    return {"workload": "unrolled shearlet/wavelet soft-threshold separation, %d x %dx%d FP32, %d iterations, "
                        "NSST [8,8,16,16] + 3-level D4, per-band learnable thresholds" % (nimg, L, L, iters),
            "forward_ms_per_iter": round(t_f * 1e3 / reps / iters, 3),
            "fwd_bwd_ms_per_iter": round(t_fb * 1e3 / reps / iters, 3)}


VOL_DIRS = (8, 8, 4)


def vol_bytes_model(sides, band_first):
    """Band-stack byte model of one 3D volume: input and output once, every
    band written once and read once (4 B per voxel each)."""
    n = sides[0]
    b = 8.0 * n ** 3
    for l in range(len(band_first) - 1):
        b += 8.0 * (band_first[l + 1] - band_first[l]) * sides[l] ** 3
    return b


def cpu_baseline_3d(n, sigma, nbands_sample=6):
    """FP64 numpy oracle of the volume denoise on one host core, extrapolated
    from a bounded sample: the band-pass decomposition and reconstruction of
    one volume plus `nbands_sample` thresholded 'same' convolutions per level,
    scaled to all bands of the level."""
    import numpy as np

    from oracle import nsst3d as o3
    from shearcuda_amd.synth import noisy_ellipsoids
    _, x = noisy_ellipsoids(n, sigma=sigma, seed=700)
    F = o3.meyer_filters_3d(list(VOL_DIRS))
    t0 = time.perf_counter()
    subs = o3.pyr_nd_dec(x, len(VOL_DIRS), 1)
    o3.pyr_nd_rec(subs, 1)
    total = time.perf_counter() - t0
    sampled = 0
    for l in range(len(VOL_DIRS)):
        bands = o3.band_list(F, l)
        conv = o3._SameConv(subs[l], F[0][l][0][0].shape)
        acc = np.zeros(subs[l].shape)
        t0 = time.perf_counter()
        for f in bands[:nbands_sample]:
            b = conv(f)
            acc += np.where(np.abs(b) <= 3 * sigma * 0.05, 0.0, b)
        total += (time.perf_counter() - t0) * len(bands) / nbands_sample
        sampled += nbands_sample
    return {"value": round(n ** 3 / total / 1e6, 5), "unit": "Mvoxel/s", "cores": 1, "kind": "port",
            "sample": "one %d^3 volume: band-pass dec+rec and %d of %d band convolutions (%d per level) timed, "
                      "extrapolated to all bands (%.1f s per volume)" % (n, sampled, 3 * sum(d * d for d in VOL_DIRS),
                                                                         nbands_sample, total)}


def bench_3d(steps=5, warmup=1, nvol=4, n=128, sigma=25.0, group=8, cpu=True, plane_cols=-1):
    """3D shearlet volume denoise of a batch of 4 x 128^3 FP32 volumes
    (8 random filled ellipsoids + N(0, 25^2) each) as one graph launch:
    Fourier band-pass, 3 levels (Pyr_mode 1), 3 cones with 8x8 / 8x8 / 4x4
    directions, hard threshold 3 sigma bandEst per band, reconstruction."""
    import numpy as np
    import torch

    from shearcuda_amd import api, shearlet3d
    from shearcuda_amd.synth import noisy_ellipsoids, psnr, white_noise_volume
    t0 = time.perf_counter()
    sh = shearlet3d.Shear3D(n, VOL_DIRS, 1)
    prep_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    est, _ = shearlet3d.nsst_scalars_3d(sh, torch.tensor(white_noise_volume(n), dtype=torch.float32, device="cuda"))
    scalars_s = time.perf_counter() - t0
    pairs = [noisy_ellipsoids(n, sigma=sigma, seed=700 + i) for i in range(nvol)]
    x = torch.tensor(np.stack([p[1] for p in pairs]), dtype=torch.float32, device="cuda")
    th = shearlet3d.thresholds_3d(sh, est, [sigma] * nvol, (3.0, 3.0, 3.0))
    plan = shearlet3d.Plan3D(sh, nvol, group=group)
    plan.set_schedule(group=group, plane_cols=plane_cols)
    out = torch.empty_like(x)
    g = api.capture(lambda: plan.denoise(x, th, out))
    el = _timed(g.launch, steps, warmup)
    ms = el * 1e3 / steps
    res = out.cpu().numpy()
    vox = nvol * n ** 3
    byt = vol_bytes_model(sh.sides, sh.band_first) * nvol
    info = {"workload": "3D shearlet volume denoise, %d x %d^3 FP32 per launch, Fourier band-pass 3 levels "
                        "(Pyr_mode 1: %s), 3 cones x %s directions (%d bands of 24^3 filters), FFT grids %s, "
                        "T=3 sigma bandEst, sigma=%g, one hipGraph launch" % (
                            nvol, n, sh.sides, "/".join("%dx%d" % (d, d) for d in VOL_DIRS), sh.nbands,
                            sh.fft_sides, sigma),
            "ms_per_launch": round(ms, 3), "Mvoxel_s": round(vox / (ms * 1e-3) / 1e6, 2),
            "bytes_model_per_launch": byt,
            "path_roofline_frac": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "psnr_in": [round(psnr(p[0], p[1]), 3) for p in pairs],
            "psnr_out": [round(psnr(p[0], res[i]), 3) for i, p in enumerate(pairs)],
            "band_group": group, "plane_cols": plane_cols, "dict_bytes": sh.nbytes(), "plan_bytes": plan.nbytes(),
            "dict_prep_s": round(prep_s, 2), "noise_scalars_s": round(scalars_s, 2)}
    if cpu:
        info["cpu_baseline"] = cpu_baseline_3d(n, sigma)
    g.close()
    plan.close()
    sh.close()
    return info


# ---- HBM traffic ledger of the plain step (rocprofv3 counters, live) -------------
LEDGER_LAUNCHES = 3
LEDGER_COPY_MIB = 256


def ledger_child(spec):
    """Child process of traffic_ledger.  Without a schedule: autotune the plain
    step and write the schedule.  With one: run the step's graph
    LEDGER_LAUNCHES times, then one device copy of LEDGER_COPY_MIB MiB (the
    counters' calibration), and write the step's stage names in launch
    order.  Run under rocprofv3 --pmc for the counted passes."""
    import torch

    from shearcuda_amd import api
    torch.cuda.set_device(0)
    L, sigma = spec["L"], spec["sigma"]
    shear, xg, sc = plain_setup(L, sigma)
    out = {}
    if spec.get("schedule") is None:
        out["schedule"] = api.autotune(shear)
    else:
        apply_schedule(shear, spec["schedule"])
        out["stages"] = [st["name"] for st in profile_plain(shear, xg, sc, sigma, 1)]
        g, _, _ = api.denoise_graph(xg, shear, sigma, sc, coeffs=False)
        # the copy's buffers are filled before the launches, so that the copy
        # is the only dispatch after the last launch
        src = torch.ones(LEDGER_COPY_MIB * (1 << 20) // 4, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(LEDGER_LAUNCHES):
            g.launch()
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        g.close()
    shear.close()
    with open(spec["out"], "w") as f:
        json.dump(out, f)


def _run_child(spec, counter=None, timeout=240):
    """bench.py --ledger-child in a fresh process (optionally under
    rocprofv3 --pmc counter); returns (result dict, counter CSV rows)."""
    d = tempfile.mkdtemp(prefix="shx_ledger_", dir="/tmp")
    spec = dict(spec, out=os.path.join(d, "child.json"))
    cmd = [sys.executable, os.path.abspath(__file__), "--ledger-child", json.dumps(spec)]
    if counter:
        cmd = ["timeout", "-s", "KILL", str(timeout), shutil.which("rocprofv3"), "--pmc", counter,
               "--output-format", "csv", "-d", d, "-o", "pmc", "--"] + cmd
    else:
        cmd = ["timeout", "-s", "KILL", str(timeout)] + cmd
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    try:
        if r.returncode != 0:
            tail = r.stdout.decode(errors="replace")[-600:]
            raise RuntimeError("ledger child (%s) exited %d: %s" % (counter or "tuning", r.returncode, tail))
        with open(spec["out"]) as f:
            res = json.load(f)
        rows = []
        if counter:
            files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
            if not files:
                raise RuntimeError("rocprofv3 wrote no counter_collection.csv")
            with open(files[0]) as f:
                rows = [row for row in csv.DictReader(f) if row.get("Counter_Name") == counter]
        return res, rows
    finally:
        shutil.rmtree(d, ignore_errors=True)


def _dispatches(rows):
    """(dispatch id, kernel name, counter value in bytes) in dispatch order;
    FETCH_SIZE and WRITE_SIZE count KiB."""
    out = {}
    for row in rows:
        k = int(row["Dispatch_Id"])
        name = row["Kernel_Name"]
        v = float(row["Counter_Value"]) * 1024.0
        out[k] = (name, out.get(k, (name, 0.0))[1] + v)
    return [(k, n, v) for k, (n, v) in sorted(out.items())]


def _split_launch(disp, nk):
    """The last graph launch's nk dispatches and the calibration copy's
    dispatches after them."""
    own = [i for i, (_, n, _) in enumerate(disp) if "shx::" in n or n.startswith("k_")]
    if not own:
        raise RuntimeError("no kernels of the library in the counter run")
    last = own[-1]
    step = disp[last - nk + 1:last + 1]
    copy = [d for d in disp[last + 1:] if "copy" in d[1].lower()]
    prev = disp[last - 2 * nk + 1:last - nk + 1]
    if len(step) != nk or not copy or [n for _, n, _ in prev] != [n for _, n, _ in step]:
        raise RuntimeError("could not separate one graph launch and the calibration copy")
    return step, copy


def traffic_ledger(L, sigma):
    """Per-kernel HBM bytes of one plain step from two rocprofv3 --pmc passes
    (FETCH_SIZE, WRITE_SIZE) of a child process running the step's graph,
    each scaled by what the same counter reads for a device copy of
    LEDGER_COPY_MIB MiB.  A tuning child first fixes the schedule, which the
    counted passes and the timed run then share.  Runs before this process
    touches the GPU (the children are separate processes).  Returns the
    ledger, or {"error": ...} when the counters are unavailable."""
    if not shutil.which("rocprofv3"):
        return {"error": "rocprofv3 not found"}
    try:
        base = {"L": L, "sigma": sigma}
        tuned, _ = _run_child(dict(base, schedule=None))
        sched = tuned["schedule"]
        spec = dict(base, schedule=sched)
        res_f, rows_f = _run_child(spec, "FETCH_SIZE")
        res_w, rows_w = _run_child(spec, "WRITE_SIZE")
        names = res_f["stages"]
        nk = len(names)
        step_f, copy_f = _split_launch(_dispatches(rows_f), nk)
        step_w, copy_w = _split_launch(_dispatches(rows_w), nk)
        cbytes = float(LEDGER_COPY_MIB << 20)
        fscale = cbytes / sum(v for _, _, v in copy_f)
        wscale = cbytes / sum(v for _, _, v in copy_w)
        kernels = []
        for stage, (_, kn, fv), (_, kw, wv) in zip(names, step_f, step_w):
            if kn != kw:
                raise RuntimeError("the two counter passes launched different kernels")
            kernels.append({"stage": stage, "kernel": kn.split("(")[0].replace("void ", ""),
                            "fetch_bytes": round(fv * fscale), "write_bytes": round(wv * wscale),
                            "hbm_bytes": round(fv * fscale + wv * wscale)})
        total = sum(k["hbm_bytes"] for k in kernels)
        model = 400.0 * L * L
        return {"schedule": sched, "kernels": kernels, "step_hbm_bytes": total, "model_bytes": model,
                "ratio_to_model": round(total / model, 3),
                "calibration": {"copy_mib": LEDGER_COPY_MIB, "fetch_scale": round(fscale, 4),
                                "write_scale": round(wscale, 4)},
                "method": "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, over a child running the step's graph %d "
                          "times; the last launch's dispatches" % LEDGER_LAUNCHES}
    except Exception as e:  # counters are a report, not the measurement
        return {"error": str(e)[-800:]}


# ---- kernel-trace windows (profiling aid) -------------------------------------------
WINDOW_GAP_S = 0.05


def trace_windows(which):
    """Set a workload up as the bench does (autotuned), then launch its timed
    unit three times between idle gaps of WINDOW_GAP_S, so that a kernel trace
    of this process separates single launches by time (summarize_windows).
    which: plain (the 512^2 image-only step), plain1024 (the same at 1024^2),
    batched (the mix in one graph), adjoint (2 unrolled iterations forward,
    then forward + backward)."""
    import torch

    from shearcuda_amd import api
    torch.cuda.set_device(0)
    units = []
    keep = []
    if which in ("plain", "plain1024"):
        L = 512 if which == "plain" else 1024
        shear, xg, sc = plain_setup(L, 20.0)
        api.autotune(shear)
        g, _, _ = api.denoise_graph(xg, shear, 20.0, sc, coeffs=False)
        units.append(("plain%d" % L, g.launch))
        keep += [shear, g]
    elif which == "batched":
        import numpy as np

        from shearcuda_amd.synth import noisy_ellipses, white_noise
        rng = np.random.default_rng(2024)
        plans, seed = [], 100
        for L, n in BATCH_MIX:
            shear = api.make_dictionary(L)
            sc = api.nsst_scalars_cuda(shear, torch.from_numpy(white_noise(L).astype(np.float32)).cuda())
            sig = rng.uniform(5.0, 40.0, size=n)
            xs = np.stack([noisy_ellipses(L, sigma=float(sig[i]), seed=seed + i)[1] for i in range(n)])
            seed += n
            b = api.Batch(shear, n)
            api.autotune(shear, b)
            plans.append(dict(b=b, x=torch.from_numpy(xs.astype(np.float32)).cuda(),
                              th=api.batch_thresholds(shear, sig, sc), out=torch.empty((n, L, L), device="cuda")))
            keep += [shear, b]
        g = api.capture(lambda: [p["b"].denoise(p["x"], None, p["th"], p["out"]) for p in plans])
        units.append(("batched_mix", g.launch))
        keep += [g, plans]
    elif which == "adjoint":
        fwd, fwd_bwd, hold = _adjoint_setup(iters=2)
        units += [("adjoint_forward_2iter", fwd), ("adjoint_fwd_bwd_2iter", fwd_bwd)]
        keep.append(hold)
    else:
        raise ValueError(which)
    for _, fn in units:   # warm: caching allocator pools, first launches
        fn()
        fn()
    torch.cuda.synchronize()
    for name, fn in units:
        for _ in range(3):
            time.sleep(WINDOW_GAP_S)
            fn()
            torch.cuda.synchronize()
            print("window", name, flush=True)
    time.sleep(WINDOW_GAP_S)


def summarize_windows(path, names):
    """Markdown tables of the launch windows of a kernel trace written under
    path (rocprofv3 --kernel-trace --output-format csv) by --trace-windows:
    windows are separated by idle gaps of at least half WINDOW_GAP_S; the last
    3 * len(names) windows are the measured launches, the middle one of each
    three is tabulated."""
    files = glob.glob(os.path.join(path, "**", "*kernel_trace.csv"), recursive=True)
    if not files:
        raise SystemExit("no kernel_trace.csv under %s" % path)
    with open(files[0]) as f:
        ks = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]) for r in csv.DictReader(f))
    wins, cur, last = [], [], None
    for k in ks:
        if last is not None and k[0] - last > WINDOW_GAP_S * 0.5e9:
            wins.append(cur)
            cur = []
        cur.append(k)
        last = k[1] if last is None else max(last, k[1])
    wins.append(cur)
    need = 3 * len(names)
    if len(wins) < need:
        raise SystemExit("found %d windows, need %d" % (len(wins), need))
    out = []
    for i, name in enumerate(names):
        w = wins[len(wins) - need + 3 * i + 1]
        t0 = min(k[0] for k in w)
        wall = (max(k[1] for k in w) - t0) / 1e3
        agg = {}
        for a, b, n in w:
            short = n.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")
            c, t = agg.get(short, (0, 0.0))
            agg[short] = (c + 1, t + (b - a) / 1e3)
        tot = sum(t for _, t in agg.values())
        out.append("### %s: one launch, %d kernels, %.1f us from first start to last end, %.1f us of kernels"
                   % (name, len(w), wall, tot))
        out.append("")
        out.append("| kernel | launches | total (us) | mean (us) | share of kernel time |")
        out.append("|---|---:|---:|---:|---:|")
        for n, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
            out.append("| `%s` | %d | %.1f | %.1f | %.1f%% |" % (n, c, t, t / c, 100.0 * t / tot))
        out.append("")
    print("\n".join(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU oracle work (0 = skip)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-plain-1024", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-3d", action="store_true")
    ap.add_argument("--only", choices=["3d", "adjoint", "batched"], default=None,
                    help="run one secondary workload alone (profiling); prints its result object")
    ap.add_argument("--vol-group", type=int, default=8, help="bands per launch of the 3D path")
    ap.add_argument("--vol-steps", type=int, default=5)
    ap.add_argument("--vol-plane-cols", type=int, default=-1,
                    help="3D directional loop: -1 separate row/column passes, 0/32/64 the fused plane kernel")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB first (HBM counter calibration)")
    ap.add_argument("--no-ledger", action="store_true", help="skip the rocprofv3 traffic ledger of the plain step")
    ap.add_argument("--ledger-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--trace-windows", choices=["plain", "plain1024", "batched", "adjoint"], default=None,
                    help="profiling aid: launch one workload's timed unit between idle gaps (run under rocprofv3)")
    ap.add_argument("--summarize-windows", default=None, metavar="DIR",
                    help="print per-launch kernel tables from a --trace-windows kernel trace under DIR")
    args = ap.parse_args()
    if args.ledger_child:
        ledger_child(json.loads(args.ledger_child))
        return
    if args.summarize_windows:
        names = {"plain": ["plain512"], "plain1024": ["plain1024"], "batched": ["batched_mix"],
                 "adjoint": ["adjoint_forward_2iter", "adjoint_fwd_bwd_2iter"]}[args.trace_windows or "plain"]
        summarize_windows(args.summarize_windows, names)
        return
    if args.trace_windows:
        trace_windows(args.trace_windows)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the NSST path has no multi-GPU form here); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    cpu = None
    if args.cpu_budget > 0 and not args.only:
        cpu = cpu_baseline(args.size, 20.0, args.cpu_budget)   # before anything touches the GPU
    ledger = None
    if not args.no_ledger and not args.only:
        ledger = traffic_ledger(args.size, 20.0)   # child processes, also before this one touches the GPU
    import torch
    torch.cuda.set_device(0)
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.only:
        r = {"3d": lambda: bench_3d(steps=args.vol_steps, group=args.vol_group, plane_cols=args.vol_plane_cols),
             "adjoint": bench_adjoint,
             "batched": lambda: bench_batched(max(5, args.steps // 20), 2)}[args.only]()
        print(json.dumps(r), flush=True)
        return
    L, sigma = args.size, 20.0
    ms, ms_coeffs, stages, sched = bench_plain(L, sigma, args.steps, args.warmup,
                                               ledger.get("schedule") if ledger else None)
    value = L * L / (ms * 1e-3) / 1e6
    dom = max(stages, key=lambda s: s["ms"])
    # HBM bytes only from this run's own counter passes (null without them)
    traffic, traffic_src = None, None
    if ledger and "kernels" in ledger and len(ledger["kernels"]) == len(stages):
        traffic = ledger["kernels"][stages.index(dom)]["hbm_bytes"]
        traffic_src = "ledger"
    roof = {"kernel": dom["name"], "bound": "hbm", "achieved": round(dom["GBs"], 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(dom["GBs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": traffic_src, "bytes_per_launch": dom["bytes"], "ms_per_launch": round(dom["ms"], 5)}
    path_bytes = 400.0 * L * L
    path_gbs = path_bytes / (ms * 1e-3) / 1e9
    res = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpixel/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain 2D NSST hard-threshold denoise (dec+thresh+rec), one %dx%d FP32 image per step, "
                               "maxflat 4 levels, directions [8,8,16,16], Meyer sizes [32,32,16,16], T=[0,3,4], "
                               "sigma=20, one hipGraph launch per step returning the denoised image" % (L, L),
                   "L": L, "images_per_step": 1, "form": "plain", "schedule": sched},
        "with_coeffs": {"ms_per_step": round(ms_coeffs, 5),
                        "Mpixel_s": round(L * L / (ms_coeffs * 1e-3) / 1e6, 2),
                        "note": "the same step also writing the 48 unthresholded band coefficient planes"},
        "roofline": roof,
        "path_roofline": {"bytes_per_step": path_bytes, "achieved": round(path_gbs, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(path_gbs / HBM_PEAK_GBS, 4)},
        "path_valu_roofline": {"flop_per_step": round(path_flops_per_px(L) * L * L),
                               "achieved": round(path_flops_per_px(L) * L * L / (ms * 1e-3) / 1e12, 2),
                               "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": valu_frac(L, L * L, ms * 1e-3)},
        "kernels": [{k: (round(v, 5) if isinstance(v, float) else v) for k, v in s.items()} for s in stages],
    }
    if ledger is not None:
        res["traffic_ledger"] = ledger
    if L == 512 and not args.no_plain_1024:
        ms1k, ms1kc, _, _ = bench_plain(1024, sigma, max(20, args.steps // 4), 5)
        res["plain_1024"] = {"ms_per_image": round(ms1k, 5), "Mpixel_s": round(1024 * 1024 / (ms1k * 1e-3) / 1e6, 2),
                             "Mpixel_s_with_coeffs": round(1024 * 1024 / (ms1kc * 1e-3) / 1e6, 2),
                             "path_roofline_frac": round(400.0 * 1024 * 1024 / (ms1k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                             "path_valu_frac": valu_frac(1024, 1024 * 1024, ms1k * 1e-3)}
        res["plain_1024"]["dictionary_build_s"] = dictionary_build_times(1024)
    if not args.no_batched:
        res["batched"] = bench_batched(max(5, args.steps // 20), 2)
    if not args.no_adjoint:
        res["adjoint"] = bench_adjoint()
        res["separate"] = bench_separate()
    if not args.no_3d:
        res["volume"] = bench_3d(steps=args.vol_steps, group=args.vol_group, cpu=args.cpu_budget > 0,
                                 plane_cols=args.vol_plane_cols)
    if cpu is not None:
        res["cpu_baseline"] = cpu
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

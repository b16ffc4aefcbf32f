#!/usr/bin/env python3
"""Splotch particle-splat hot path on one MI355X.

Workload (default): 16M particles, 2048 x 2048 RGB float32 image — 8M gas
(type 0, depth-sorted emission-absorption, brightness 0.5) + 8M stars
(type 1, additive, hsml 0.05, brightness 2.0), seeded synthetic Plummer-halo
fields, camera (50,50,-150) -> (50,50,50), sky (0,1,0), fov 30.

A step is one frame of the full hot path with the particles already resident
in HBM: project + colorize, depth sort, binning, splat of both passes.  The
timed region is K frames bracketed by device synchronisation.

The JSON line also carries:
  roofline      for the dominant kernel (k_splat8 of the ordered pass, the
                8x8-block splat of the normal work items), algorithmic bytes
                per launch = 24 B per active ordered particle + 12 B per pixel,
                over that launch's duration from HIP events the library
                records around it on its own stream (average of the profiled
                frames); traffic from rocprofv3 FETCH_SIZE + WRITE_SIZE (KiB
                per dispatch, scaled by a known-size copy kernel's counts);
  valu          the limit the splat actually meets: VALU instructions of the
                ordered splat per frame (rocprofv3 SQ_INSTS_VALU, committed
                summary) over the ordered splat stage's live duration, against
                the issue peak (256 CUs x 4 SIMDs x one wave64 instruction per
                4 clocks at 2.4 GHz);
  lds           the LDS read peak measured by a microbenchmark on this GPU
                (libmembench), and the splat kernels' LDS bytes per dispatch
                (rocprofv3 LDS counters, scaled by the microbenchmark's own
                counts; committed summary) over their event durations;
  parity        per workload, the max per-channel relative error of the GPU
                image (denominator max(|ref|, 1e-6 max|ref|)) against the
                oracle (stable depth order), the genuine reference renderer
                (std::sort depth order, where it differs) and the FP64 oracle
                (true exp, double state and accumulation, same footprints),
                and the depth ties among the ordered particles;
  c0            the plain workload: 1M particles, 800^2, additive;
  batched       the batched workload (64 views, synth.batch_jobs): one
                splhip_render_batched call, a replay of its captured HIP graph,
                the autotuned batch, and one splhip_render per view;
  fwd_bwd       the differentiable workload: 2M particles, 512^2, additive,
                and a 500k ordered variant; loss sum (pic - target)^2 against
                the render of a perturbed copy; forward vs forward + backward;
  epilogue, normalize, ingest   the steps either side of the path;
  fragments     the frame's per-pixel updates (oracle count, committed
                summary) over the frame time and over the ordered splat stage;
  hbm_copy      a hand-written device copy (1 GiB) beside the 8 TB/s datasheet;
  cpu_baseline  the genuine reference host renderer (oracle/_ref, built from
                the Splotch sources; the oracle restatement if that library is
                absent) on this host's cores: the full frame, one warm-up run
                then the median of three.  c0, batched (per view, summed) and
                fwd_bwd (forward) carry their own, timed the same way.
"""
from __future__ import annotations

import argparse
import contextlib
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# A frame's passes and long lists run on up to 8 HIP streams (the caller's, a
# second pass stream and two sets of three auxiliary streams).  HIP maps a
# process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default), and
# streams that share a queue run their work in turn.  Set here, before the
# HIP runtime starts (torch is imported later).
if int(os.environ.get("GPU_MAX_HW_QUEUES") or 4) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

METRIC = "ms/frame & Mparticles/s (16M, 2048²); batched and fwd+bwd vs plain"
HBM_PEAK_GBS = 8000.0
VALU_PEAK = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions per second (one per SIMD per 4 clocks)
LDS_SPEC_GBS = 256 * 128 * 2.4e9 / 1e9  # 128 B per clock per CU (32 banks x 4 B)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


@contextlib.contextmanager
def quiet_stdout():
    """Silence C-level stdout (the reference prints progress lines)."""
    sys.stdout.flush()
    fd = os.dup(1)
    nul = os.open(os.devnull, os.O_WRONLY)
    os.dup2(nul, 1)
    try:
        yield
    finally:
        os.dup2(fd, 1)
        os.close(fd)
        os.close(nul)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    return O


# ------------------------------------------------------------------ CPU side

def median_run(fn, reps=3):
    """One warm-up call, then the median wall time of `reps` calls; returns
    (seconds, times, the last call's result)."""
    fn()
    times, out = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
    return sorted(times)[len(times) // 2], times, out


def cpu_render(O, field, view, pal, aeq, br=1.0):
    """The reference host renderer (host_rendering, genuine build) on one
    type's particles, or the oracle restatement (std::sort order) without it."""
    import numpy as np
    if O.ref_available():
        params = {"xres": view.xres, "yres": view.yres, "fov": view.fov, "a_eq_e": aeq,
                  "brightness0": br}
        with quiet_stdout():
            return O.ref_host_rendering(field, view, [pal], params)
    pic, _ = O.host_rendering(field, view, [pal], a_eq_e=aeq, brightness=[br], stable=False)
    return pic.astype(np.float32)


def type_field(synth, f, t):
    import numpy as np
    sel = f.type == t
    return synth.Field(f.x[sel], f.y[sel], f.z[sel], f.hsml[sel], f.inten[sel], f.col[sel],
                       np.zeros(int(sel.sum()), np.uint8))


def cpu_two_pass(O, synth, field, view, pal):
    """The reference renderer on the gas (ordered, brightness 0.5) and the
    stars (additive, brightness 2.0) of a two-type field: image P_gas + P_stars."""
    import numpy as np
    gas = type_field(synth, field, 0)
    stars = type_field(synth, field, 1)

    def run():
        a = cpu_render(O, gas, view, pal, False, 0.5)
        b = cpu_render(O, stars, view, pal, True, 2.0)
        return (a + b).astype(np.float32)
    return run


def cpu_info(O, value, unit, sample, seconds, times):
    threads = int(O.lib().orc_num_threads())
    return {"value": value, "unit": unit, "cores": threads,
            "kind": "reference" if O.ref_available() else "port",
            "sample": "%s; %s host renderer with OpenMP on %d threads (%s); one warm-up run, "
                      "median of %d runs (%s s)"
                      % (sample, "genuine Splotch" if O.ref_available() else "oracle restatement",
                         threads, cpu_model(), len(times),
                         ", ".join("%.3f" % t for t in times)),
            "seconds": round(seconds, 4)}


def max_rel(a, b):
    """Max per-channel relative error |a - b| / max(|b|, 1e-6 max|b|)."""
    import numpy as np
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    den = np.maximum(np.abs(b), 1e-6 * max(float(np.abs(b).max()), 1e-300))
    return float((np.abs(a - b) / den).max())


def max_abs_over_max(a, b):
    """max |a - b| / max |b|: the error at the image's own scale."""
    import numpy as np
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return float(np.abs(a - b).max() / max(float(np.abs(b).max()), 1e-300))


def oracle_images(O, field, view, pals, modes, brightness, fp64):
    """The oracle's image of a field whose type t takes part in the passes of
    modes[t] (1 additive, 0 ordered, 2 both), the ordered pass in stable (z,
    index) order: float32 (render) or, with fp64, the FP64 twin (render64:
    true exp, double state from project64, double accumulation, on the float
    footprints).  Also returns the projected particles."""
    import numpy as np
    p = O.project_colorize(field, view, pals, brightness=brightness)
    typ = p.type if field.type is not None else np.zeros(p.n, np.uint8)
    ordm = np.isin(typ, [t for t, m in enumerate(modes) if m != 1]).astype(np.uint8)
    addm = np.isin(typ, [t for t, m in enumerate(modes) if m != 0]).astype(np.uint8)
    s64 = O.project64(field, view, pals, brightness=brightness) if fp64 else None
    pic = None
    if ordm.any():
        g = p.take(np.arange(p.n))
        g.active = p.active * ordm
        perm = O.sort_perm(g, 1, stable=True)
        gs = g.take(perm)
        if fp64:
            pic = O.render64(gs, {k: v[perm] for k, v in s64.items()}, view.xres, view.yres,
                             False)
        else:
            pic = O.render(gs, view.xres, view.yres, False)
    if addm.any():
        a = p.take(np.arange(p.n))
        a.active = p.active * addm
        img = O.render64(a, s64, view.xres, view.yres, True) if fp64 else \
            O.render(a, view.xres, view.yres, True)
        pic = img if pic is None else pic + img
    return (pic.astype(np.float64) if fp64 else pic.astype(np.float32)), p


def z_ties(p, ordered_mask):
    """Equal depth keys among the ordered pass's active particles (std::sort
    may order these differently from the stable (z, index) order)."""
    import numpy as np
    z = np.sort(p.z[(p.active != 0) & ordered_mask])
    _, counts = np.unique(z, return_counts=True)
    return {"active_ordered": int(z.size), "tied_pairs": int((z[1:] == z[:-1]).sum()),
            "particles_in_ties": int(counts[counts > 1].sum()),
            "largest_tie": int(counts.max()) if counts.size else 0}


# ------------------------------------------------------------------ GPU side

def gpu_time(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def membench():
    lib = ctypes.CDLL(os.path.join(ROOT, "splotch_amd", "libmembench.so"))
    lib.lds_peak_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)]
    lib.hbm_copy_run.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_float)]
    return lib


def lds_peak(grid=2048, iters=12000):
    """LDS read bandwidth of conflict-free ds_read_b32 / b64 / b128 streams;
    the best is the measured peak."""
    runs = {}
    for words, name in ((1, "b32"), (2, "b64"), (4, "b128")):
        ms, nbytes = ctypes.c_float(), ctypes.c_double()
        it = iters * 4 // words  # about the same time per variant
        rc = membench().lds_peak_run(0, grid, it, words, ctypes.byref(ms), ctypes.byref(nbytes))
        if rc != 0:
            raise RuntimeError("lds_peak_run failed (%d)" % rc)
        runs[name] = {"ms": round(ms.value, 4), "bytes": nbytes.value,
                      "GBs": round(nbytes.value / (ms.value * 1e-3) / 1e9, 1)}
    best = max(runs, key=lambda k: runs[k]["GBs"])
    return {"kernel": "k_lds_read<words> (conflict-free ds_read_%s)" % best,
            "GBs": runs[best]["GBs"], "variants": runs}


def hbm_copy(nbytes=1 << 30, reps=10):
    ms = ctypes.c_float()
    rc = membench().hbm_copy_run(0, nbytes, reps, 0, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError("hbm_copy_run failed (%d)" % rc)
    return {"kernel": "k_copy (16-byte vectors, 8 workgroups per CU)", "bytes_moved": 2 * nbytes,
            "ms": round(ms.value, 4), "GBs": round(2 * nbytes / (ms.value * 1e-3) / 1e9, 1),
            "datasheet_GBs": HBM_PEAK_GBS}


def profile_value(name, key):
    """A `key: value` line of a committed rocprofv3 summary under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def bench_c0(S, R, torch, O, synth, args):
    """The plain workload: 1M particles (seed 1000), 800^2, additive."""
    fld = synth.plummer_field(1_000_000, 1000)
    view = synth.default_view(800)
    pal = synth.default_palette()
    dev = S.to_device(fld)
    cam = S.camera_setup(view)
    prm = S.params(a_eq_e=True)
    pic = torch.empty(800 * 800 * 3, dtype=torch.float32, device="cuda")
    ms = gpu_time(torch, lambda: R.render(dev, cam, prm, pic), 20, 2)
    R.set_profiling(True)
    R.render(dev, cam, prm, pic)
    stages, counts = R.stage_times()
    R.set_profiling(False)
    got = pic.cpu().numpy().reshape(800, 800, 3)
    out = {"workload": "plain_1M_800sq", "particles": fld.n, "image": [800, 800],
           "ms_per_frame": round(ms, 4), "mparticles_s": round(fld.n / (ms * 1e-3) / 1e6, 2),
           "stages_ms": {k: round(v, 4) for k, v in stages.items() if v},
           "counts": counts}
    want, _ = oracle_images(O, fld, view, [pal], [1], None, False)
    f64, _ = oracle_images(O, fld, view, [pal], [1], None, True)
    par = {"vs_oracle": max_rel(got, want), "vs_fp64": max_rel(got, f64),
           "vs_fp64_abs_over_max": max_abs_over_max(got, f64)}
    if not args.no_cpu:
        sec, times, ref = median_run(lambda: cpu_render(O, fld, view, pal, True))
        out["cpu_baseline"] = cpu_info(O, round(fld.n / sec / 1e6, 4), "Mparticles/s",
                                       "the full 1M / 800^2 frame", sec, times)
        out["cpu_baseline"]["speedup"] = round(sec * 1e3 / ms, 1)
        par["vs_reference"] = max_rel(got, ref)
    out["parity"] = par
    log("c0: %.3f ms/frame, parity %s" % (ms, par))
    return out


def bench_batched(S, R, torch, O, synth, args, n_jobs=64):
    """Batched workload: n_jobs views (synth.batch_jobs) as one batch, as a
    replay of the batch's captured HIP graph, autotuned, and one
    splhip_render call per view.  Aggregate Mparticles/s of each."""
    import numpy as np
    steps, warmup = max(2, args.steps // 2), 1
    jobs = synth.batch_jobs(n_jobs)
    views = []
    for bj in jobs:
        views.append((S.to_device(bj.field), S.camera_setup(bj.view), S.params(a_eq_e=bj.additive),
                      torch.empty(bj.view.xres * bj.view.yres * 3, dtype=torch.float32,
                                  device="cuda")))
    ntot = sum(bj.field.n for bj in jobs)
    res = {}
    res["plain"] = gpu_time(torch, lambda: [R.render(*v) for v in views], steps, warmup)
    res["batched"] = gpu_time(torch, lambda: R.render_batched(views), steps, warmup)
    ref = [v[3].cpu().numpy() for v in views]
    # default tuning
    default = R.tuning()
    # the batch captured into a HIP graph and replayed
    R.capture_reserve(views, 1.25)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            R.render_batched(views)
    torch.cuda.synchronize()
    res["replay"] = gpu_time(torch, g.replay, steps, warmup)
    overflow = R.capture_status()
    same = all(np.array_equal(v[3].cpu().numpy().view(np.uint32), r.view(np.uint32))
               for v, r in zip(views, ref))
    del g
    tuned = [(d, c, S.params(a_eq_e=bj.additive, hip_autotune=1), p)
             for (d, c, _, p), bj in zip(views, jobs)]
    res["autotuned"] = gpu_time(torch, lambda: R.render_batched(tuned), steps, warmup)
    chosen = R.tuning()
    mps = {k: ntot / (v * 1e-3) / 1e6 for k, v in res.items()}
    log("batched %d views, %d particles: %s ms (%s)" % (n_jobs, ntot, res, chosen))
    out = {"views": n_jobs, "particles": ntot, "sides": sorted({bj.view.xres for bj in jobs}),
           "ms_per_batch": round(res["batched"], 3), "mparticles_s": round(mps["batched"], 3),
           "plain_ms_per_batch": round(res["plain"], 3),
           "plain_mparticles_s": round(mps["plain"], 3),
           "batched_vs_plain": round(mps["batched"] / mps["plain"], 3),
           "replay_ms_per_batch": round(res["replay"], 3),
           "replay_mparticles_s": round(mps["replay"], 3),
           "replay_vs_plain": round(mps["replay"] / mps["plain"], 3),
           "replay_images_equal_batched": bool(same), "replay_overflow_bits": overflow,
           "autotuned_ms_per_batch": round(res["autotuned"], 3),
           "autotuned_vs_plain": round(mps["autotuned"] / mps["plain"], 3),
           "autotuned_vs_default": round(res["batched"] / res["autotuned"], 3),
           "default_choice": default, "autotuned_choice": chosen}
    if not args.no_cpu:
        pal = synth.default_palette()
        per = []
        for bj in jobs:  # one warm-up pass over all views, then three; median per view
            cpu_render(O, bj.field, bj.view, pal, bj.additive)
        runs = []
        for _ in range(3):
            t = []
            for bj in jobs:
                t0 = time.perf_counter()
                cpu_render(O, bj.field, bj.view, pal, bj.additive)
                t.append(time.perf_counter() - t0)
            runs.append(t)
        per = [sorted(r[j] for r in runs)[1] for j in range(n_jobs)]
        sec = sum(per)
        out["cpu_baseline"] = cpu_info(O, round(ntot / sec / 1e6, 4), "Mparticles/s",
                                       "all %d views, each timed alone and summed" % n_jobs, sec,
                                       [sum(r) for r in runs])
        out["cpu_baseline"]["speedup_batched"] = round(sec * 1e3 / res["batched"], 1)
        out["cpu_baseline"]["speedup_replay"] = round(sec * 1e3 / res["replay"], 1)
    return out


def bench_backward(S, R, torch, O, synth, args):
    """Differentiable workload: forward, and forward + backward of
    L = sum (pic - target)^2, target = render of the same field with
    x += N(0, 0.05) per axis and I *= U(0.8, 1.2) (synth.perturbed)."""
    steps = max(3, args.steps // 2)
    out = {}
    for name, n, seed, aeq in (("additive_2M", 2_000_000, 1300, True),
                               ("ordered_500k", 500_000, 1301, False)):
        fld = synth.plummer_field(n, seed)
        tf = synth.perturbed(fld, 1302)
        view = synth.default_view(512)
        pal = synth.default_palette()
        cam = S.camera_setup(view)
        prm = S.params(a_eq_e=aeq)
        dev = S.to_device(fld)
        pic = torch.empty(512 * 512 * 3, dtype=torch.float32, device="cuda")
        target = torch.empty_like(pic)
        R.render(S.to_device(tf), cam, prm, target)

        def fwd():
            R.render(dev, cam, prm, pic)

        def fwd_bwd():
            R.render(dev, cam, prm, pic)
            return R.backward((2.0 * (pic - target)).contiguous(), n)

        ms_f = gpu_time(torch, fwd, steps, 1)
        ms_fb = gpu_time(torch, fwd_bwd, steps, 1)
        fwd()
        got = pic.cpu().numpy().reshape(512, 512, 3)
        f64, _ = oracle_images(O, fld, view, [pal], [1 if aeq else 0], None, True)
        rec = {"particles": n, "image": [512, 512], "additive": aeq, "ms_fwd": round(ms_f, 3),
               "ms_fwd_bwd": round(ms_fb, 3), "fwd_bwd_over_fwd": round(ms_fb / ms_f, 3),
               "parity_fwd_vs_fp64": max_rel(got, f64),
               "parity_fwd_vs_fp64_abs_over_max": max_abs_over_max(got, f64)}
        if not args.no_cpu:
            sec, times, ref = median_run(lambda: cpu_render(O, fld, view, pal, aeq))
            rec["cpu_baseline_fwd"] = cpu_info(O, round(n / sec / 1e6, 4), "Mparticles/s",
                                               "the forward render (the reference has no "
                                               "backward)", sec, times)
            rec["cpu_baseline_fwd"]["speedup_fwd"] = round(sec * 1e3 / ms_f, 1)
            rec["parity_fwd_vs_reference"] = max_rel(got, ref)
        log("fwd+bwd %s: %s" % (name, rec))
        out[name] = rec
    return out


def bench_epilogue(S, R, torch, pic, W, H, reps=20):
    q = S.post_from_par("")  # a_eq_e: -expm1, no adjustment, no colorbar
    rgb8 = torch.empty(3 * W * H, dtype=torch.uint8, device="cuda")
    for _ in range(2):
        R.postprocess(pic, W, H, q, rgb8=rgb8, layout=1)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        R.postprocess(pic, W, H, q, rgb8=rgb8, layout=1)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 15 * W * H

    def d2h(t):
        h = torch.empty(t.numel(), dtype=t.dtype, pin_memory=True)
        h.copy_(t)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            h.copy_(t)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / 5

    return {"kernel": "k_post_tga", "ms": round(ms, 5), "bytes": nbytes,
            "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1), "peak_GBs": HBM_PEAK_GBS,
            "frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "d2h_ms_rgb8": round(d2h(rgb8), 4), "d2h_ms_float": round(d2h(pic), 4)}


def bench_normalize(S, R, torch, dev, reps=10):
    d = {k: dev[k].clone() for k in ("inten", "col", "hsml", "type")}
    n = int(d["inten"].numel())
    q = S.norm_from_par("ptypes=2\n")
    R.normalize(d, q)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        R.normalize(d, q, ranges=False)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 30 * n
    return {"kernels": "k_norm_reduce + k_norm_clamp", "particles": n, "ms": round(ms, 4),
            "bytes": nbytes, "achieved_GBs": round(nbytes / (ms * 1e-3) / 1e9, 1),
            "peak_GBs": HBM_PEAK_GBS, "frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def bench_ingest(S, R, torch, field, reps=3):
    import tempfile
    import numpy as np
    cols = [field.x, field.y, field.z, field.hsml, field.inten, field.col]
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "particles.bin")
        with open(path, "wb") as fh:
            for c in cols:
                np.ascontiguousarray(c, ">f4").tofile(fh)
        text = "infile=%s\nnum_blocks=6\nx=1\ny=2\nz=3\nr=4\nI=5\nC1=6\n" % path
        R.read_bin_block(text)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            R.read_bin_block(text)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        nbytes = os.path.getsize(path)
    return {"particles": field.n, "file_bytes": nbytes, "ms": round(ms, 2),
            "GBs": round(nbytes / (ms * 1e-3) / 1e9, 2),
            "note": "host file read (page cache) + PCIe copy + byte swap on the GPU"}


def bench_large_share(S, torch, args):
    """200M particles (synth.large_share), 4096^2: an additive pass over all
    particles plus the ordered pass over the gas (the halo half): plain
    frames, then the 16 frames of an orbit (azimuth k * 22.5 deg, elevation 0,
    radius 150, sky +z) as one splhip_render_batched call.  CPU baseline: the
    reference host renderer on a random 1/100 subsample rendered at 4096^2,
    extrapolated to the full frame by the fragment count (labelled)."""
    import numpy as np
    from splotch_amd import synth
    O = oracle()
    t0 = time.time()
    field = synth.large_share()
    log("generated %d particles in %.1f s" % (field.n, time.time() - t0))
    view = synth.default_view(4096)
    pal = synth.default_palette()
    modes = synth.LARGE_SHARE_MODES
    dev = S.to_device(field, "cuda")
    R = S.Renderer(0)
    R.set_palette(0, *pal)
    R.set_palette(1, *pal)
    prm = S.params(ptypes=2, a_eq_e=modes, brightness=[0.5, 2.0])
    cam = S.camera_setup(view)
    pic = torch.empty(view.xres * view.yres * 3, dtype=torch.float32, device="cuda")
    for _ in range(args.warmup):
        R.render(dev, cam, prm, pic)
        torch.cuda.synchronize()
        log("large share: warm-up frame done")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        R.render(dev, cam, prm, pic)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    log("large share: timed frames done")
    R.set_profiling(True)
    R.render(dev, cam, prm, pic)
    stages, counts = R.stage_times()
    R.set_profiling(False)
    log("large share: %.1f ms/frame; stages %s; counts %s" % (ms, stages, counts))
    orbit = []
    for k in range(16):
        v = synth.orbit_view(math.radians(22.5 * k), 0.0, 150.0, 4096)
        orbit.append((dev, S.camera_setup(v), prm,
                      torch.empty(4096 * 4096 * 3, dtype=torch.float32, device="cuda")
                      if k < 2 else None))
    # the 16 frames as one batch call (each 200M-particle view is its own
    # pipeline group, rendered in chunks); two image buffers alternate
    views = [(d, c, p, orbit[k % 2][3]) for k, (d, c, p, _) in enumerate(orbit)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    R.render_batched(views)
    torch.cuda.synchronize()
    ms_orbit = (time.perf_counter() - t0) * 1e3 / 16
    log("large share: orbit done, %.1f ms/frame" % ms_orbit)
    out = {"metric": METRIC, "value": round(field.n / (ms * 1e-3) / 1e6, 3),
           "unit": "Mparticles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "large_share_200M_4096sq", "particles": field.n,
                      "gas_ordered_and_additive": int((field.type == 0).sum()),
                      "additive_only": int((field.type == 1).sum()), "image": [4096, 4096],
                      "a_eq_e_per_type": modes, "tile_size": 50,
                      "parallelism": "single GPU"},
           "stages_ms": {k: round(v, 3) for k, v in stages.items()}, "counts": counts,
           "orbit16_batched_ms_per_frame": round(ms_orbit, 3),
           "orbit16_batched_mparticles_s": round(field.n / (ms_orbit * 1e-3) / 1e6, 3)}
    R.close()
    del dev, views, orbit
    torch.cuda.empty_cache()
    idx = np.sort(np.random.default_rng(7).choice(field.n, size=field.n // 100, replace=False))
    sub = synth.Field(field.x[idx], field.y[idx], field.z[idx], field.hsml[idx],
                      field.inten[idx], field.col[idx], field.type[idx])
    # footprint sizes on the subsample: diameter 2 * 0.75 r of active particles
    p = O.project_colorize(sub, view, [pal, pal], brightness=[0.5, 2.0])
    act = p.active != 0
    diam = 1.5 * p.r[act]
    out["footprints"] = {"sample": "the 1/100 subsample, active particles",
                         "share_diameter_over_64px": round(float((diam > 64).mean()), 5),
                         "share_diameter_over_128px": round(float((diam > 128).mean()), 5),
                         "median_diameter_px": round(float(np.median(diam)), 3)}
    fr = {}
    for name, t in (("ordered", [0]), ("additive", [0, 1])):
        q = p.take(np.arange(p.n))
        q.active = (p.active * np.isin(p.type, t)).astype(np.uint8)
        fr[name] = O.fragments(q, 4096, 4096, 50)
    out["fragments_subsample"] = fr
    if not args.no_cpu:
        gas = type_field(synth, sub, 0)

        def run():
            a = cpu_render(O, gas, view, pal, False, 0.5)
            # additive over all particles, per-type brightness
            if O.ref_available():
                params = {"xres": 4096, "yres": 4096, "fov": view.fov, "a_eq_e": True,
                          "ptypes": 2, "brightness0": 0.5, "brightness1": 2.0}
                with quiet_stdout():
                    b = O.ref_host_rendering(sub, view, [pal, pal], params)
            else:
                b, _ = O.host_rendering(sub, view, [pal, pal], a_eq_e=True,
                                        brightness=[0.5, 2.0], stable=False)
            return a + b
        sec, times, _ = median_run(run)
        f_sub = fr["ordered"] + fr["additive"]
        f_full = 100 * f_sub  # a uniform 1/100 sample holds 1/100 of the fragments in expectation
        sec_full = sec * f_full / f_sub
        out["cpu_baseline"] = cpu_info(
            O, round(field.n / sec_full / 1e6, 4), "Mparticles/s",
            "EXTRAPOLATED: random 1/100 subsample (%d particles, both passes) rendered at "
            "4096^2 in %.2f s, scaled by the fragment count (%d in the sample, %d estimated "
            "for the frame)" % (sub.n, sec, f_sub, f_full), sec, times)
        out["cpu_baseline"]["speedup"] = round(sec_full * 1e3 / ms, 1)
    print(json.dumps(out), flush=True)


def count_fragments(args):
    """Fragments of the default frame per pass, from the oracle (CPU)."""
    import numpy as np
    O = oracle()
    from splotch_amd import synth
    field = synth.gas_and_stars(args.n_gas, args.n_stars, 1100, 1101)
    view = synth.default_view(args.res)
    pal = synth.default_palette()
    p = O.project_colorize(field, view, [pal, pal], brightness=[0.5, 2.0])
    out = {"workload": "gas_stars_%dM_%d" % (field.n // 1000000, args.res)}
    for name, t in (("ordered", 0), ("additive", 1)):
        q = p.take(np.arange(p.n))
        q.active = (p.active * (p.type == t)).astype(np.uint8)
        out["fragments_" + name] = O.fragments(q, view.xres, view.yres, 50)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-gas", type=int, default=8_000_000)
    ap.add_argument("--n-stars", type=int, default=8_000_000)
    ap.add_argument("--res", type=int, default=2048)
    ap.add_argument("--region", type=int, default=0, help="hip_region (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-parity", action="store_true", help="skip the oracle comparisons")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched workload")
    ap.add_argument("--no-backward", action="store_true", help="skip the fwd+bwd workload")
    ap.add_argument("--no-c0", action="store_true", help="skip the 1M / 800^2 workload")
    ap.add_argument("--workload", default="gas_stars",
                    choices=["gas_stars", "large_share", "plain_1M"],
                    help="gas_stars: 16M at 2048^2 (default); large_share: 200M at 4096^2; "
                         "plain_1M: only the 1M / 800^2 additive workload (its JSON)")
    ap.add_argument("--count-fragments", action="store_true",
                    help="print the default frame's fragments per pass (oracle, CPU) and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB hand-written copy kernel (profiler counter "
                         "calibration: 1 GiB read and 1 GiB written per dispatch)")
    ap.add_argument("--lds-peak", action="store_true",
                    help="only run the LDS read microbenchmark and print its JSON")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library renders on one device per "
              "context and has no multi-GPU path (an additive frame would shard by "
              "particles and sum the images); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    if args.count_fragments:
        count_fragments(args)
        return
    if args.copy_calibration:
        print(json.dumps(hbm_copy(reps=20)), flush=True)
        return
    if args.lds_peak:
        print(json.dumps(lds_peak()), flush=True)
        return

    import numpy as np
    import torch

    if args.workload == "large_share":
        import splotch_amd as S
        bench_large_share(S, torch, args)
        return
    if args.workload == "plain_1M":
        import splotch_amd as S
        from splotch_amd import synth
        R = S.Renderer(0)
        R.set_palette(0, *synth.default_palette())
        print(json.dumps(bench_c0(S, R, torch, oracle(), synth, args)), flush=True)
        R.close()
        return

    import splotch_amd as S
    from splotch_amd import synth
    O = oracle()

    t0 = time.time()
    field = synth.gas_and_stars(args.n_gas, args.n_stars, 1100, 1101)
    view = synth.default_view(args.res)
    pal = synth.default_palette()
    log("generated %d particles in %.1f s" % (field.n, time.time() - t0))

    dev = S.to_device(field, "cuda")
    R = S.Renderer(0)
    R.set_palette(0, *pal)
    R.set_palette(1, *pal)
    prm = S.params(ptypes=2, a_eq_e=[0, 1], brightness=[0.5, 2.0], hip_region=args.region)
    cam = S.camera_setup(view)
    pic = torch.empty(view.xres * view.yres * 3, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        R.render(dev, cam, prm, pic)
    torch.cuda.synchronize()

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        R.render(dev, cam, prm, pic)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_frame = wall * 1e3 / args.steps
    ev_ms = ev0.elapsed_time(ev1) / args.steps

    # stage breakdown: HIP events recorded by the library on the render stream,
    # averaged over a few profiled frames after the timed region
    R.set_profiling(True)
    acc = None
    nprof = 3
    for _ in range(nprof):
        R.render(dev, cam, prm, pic)
        st, counts = R.stage_times()
        acc = st if acc is None else {k: acc[k] + st[k] for k in st}
    R.set_profiling(False)
    stages = {k: round(v / nprof, 4) for k, v in acc.items()}
    torch.cuda.synchronize()
    image = pic.cpu().numpy().reshape(view.xres, view.yres, 3)

    W, H = view.xres, view.yres
    default = (args.n_gas, args.n_stars, args.res) == (8_000_000, 8_000_000, 2048)
    splat_ms = stages["ordered_k_splat8"]
    splat_bytes = 24 * counts["ordered_active"] + 12 * W * H
    achieved = splat_bytes / (splat_ms * 1e-3) / 1e9
    traffic, traffic_src = None, None
    tv = profile_value("pmc_traffic_gas_stars_16M_2048.md",
                       "traffic_bytes_per_dispatch k_splat8<0, 4>") if default else None
    if tv is not None:
        traffic = int(tv)
        traffic_src = "profiles/pmc_traffic_gas_stars_16M_2048.md"
    valu = None
    ninst = profile_value("pmc_valu_gas_stars_16M_2048.md",
                          "valu_instructions_per_frame ordered_splat") if default else None
    if ninst is not None:
        rate = ninst / (stages["ordered_splat"] * 1e-3)
        valu = {"kernels": "k_splat8<ordered> + k_splat_pc<ordered> + k_stage_ops<ordered> + "
                           "k_stage_chain<ordered> (concurrent)",
                "instructions": int(ninst), "stage_ms": stages["ordered_splat"],
                "achieved": round(rate / 1e9, 2), "peak": VALU_PEAK / 1e9,
                "unit": "G wave64 instr/s", "frac": round(rate / VALU_PEAK, 4),
                "source": "profiles/pmc_valu_gas_stars_16M_2048.md"}
    frame_bytes = 25 * field.n + 12 * W * H
    log("frame %.3f ms (events %.3f ms); stages %s; counts %s" % (ms_frame, ev_ms, stages,
                                                                  counts))

    out = {
        "metric": METRIC,
        "value": round(field.n / (ms_frame * 1e-3) / 1e6, 3),
        "unit": "Mparticles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_frame, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "gas_stars_16M_2048sq", "particles": field.n,
                   "gas_ordered": args.n_gas, "stars_additive": args.n_stars,
                   "image": [W, H], "tile_size": prm.tile_size,
                   "hip_region": prm.hip_region or 8, "parallelism": "single GPU",
                   "hip_hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"])},
        "ms_per_frame": round(ms_frame, 4),
        "ms_per_frame_events": round(ev_ms, 4),
        "stages_ms": stages,
        "counts": counts,
        "frame_hbm_fraction": round(frame_bytes / (ms_frame * 1e-3) / 1e9 / HBM_PEAK_GBS, 6),
        "roofline": {"bound": "hbm", "kernel": "k_splat8<ordered> (normal work items)",
                     "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "bytes_per_launch": splat_bytes, "launch_ms": splat_ms,
                     "note": "the splat is VALU-issue-bound (see valu); HBM is not its "
                             "limit"},
        "valu": valu,
    }
    # LDS: measured peak; the splat kernels' LDS bytes from the committed profile
    lp = lds_peak()
    lds = {"peak_measured_GBs": lp["GBs"], "peak_spec_GBs": LDS_SPEC_GBS,
           "microbenchmark": lp}
    for kname, stage in (("k_splat8<0, 4>", "ordered_k_splat8"),
                         ("k_splat_pc<0>", "ordered_k_splat_heavy")):
        b = profile_value("pmc_lds_gas_stars_16M_2048.md",
                          "lds_bytes_per_dispatch " + kname) if default else None
        if b is not None and stages.get(stage):
            gbs = b / (stages[stage] * 1e-3) / 1e9
            lds[kname] = {"bytes_per_dispatch": int(b), "ms": stages[stage],
                          "achieved_GBs": round(gbs, 1),
                          "frac_of_measured_peak": round(gbs / lp["GBs"], 4),
                          "source": "profiles/pmc_lds_gas_stars_16M_2048.md"}
    out["lds"] = lds
    fo = profile_value("fragments_gas_stars_16M_2048.md", "fragments_ordered") if default else None
    fa = profile_value("fragments_gas_stars_16M_2048.md", "fragments_additive") if default else None
    if fo is not None and fa is not None:
        out["fragments"] = {
            "ordered": int(fo), "additive": int(fa),
            "per_s_frame": round((fo + fa) / (ms_frame * 1e-3), 1),
            "ordered_per_s_splat_stage": round(fo / (stages["ordered_splat"] * 1e-3), 1),
            "source": "profiles/fragments_gas_stars_16M_2048.md"}
    out["hbm_copy"] = hbm_copy()
    out["epilogue"] = bench_epilogue(S, R, torch, pic, W, H)
    out["normalize"] = bench_normalize(S, R, torch, dev)
    out["ingest"] = bench_ingest(S, R, torch, field)
    del dev
    torch.cuda.empty_cache()
    if not args.no_c0:
        out["c0"] = bench_c0(S, R, torch, O, synth, args)
    if not args.no_batched:
        out["batched"] = bench_batched(S, R, torch, O, synth, args)
    if not args.no_backward:
        out["fwd_bwd"] = bench_backward(S, R, torch, O, synth, args)
    R.close()
    torch.cuda.empty_cache()
    ref = None
    if not args.no_cpu:
        sec, times, ref = median_run(cpu_two_pass(O, synth, field, view, pal))
        cb = cpu_info(O, round(field.n / sec / 1e6, 4), "Mparticles/s",
                      "the full 16M / 2048^2 frame (8M gas ordered with std::sort + 8M stars "
                      "additive), same params", sec, times)
        cb["speedup"] = round(out["value"] / cb["value"], 2)
        out["cpu_baseline"] = cb
    if not args.no_parity:
        par = {"tolerance": "max per-channel relative error, denominator max(|ref|, "
                            "1e-6 max|ref|)"}
        want, p = oracle_images(O, field, view, [pal, pal], [0, 1], [0.5, 2.0], False)
        par["vs_oracle_stable"] = max_rel(image, want)
        par["bitexact_vs_oracle_stable"] = bool(np.array_equal(image.view(np.uint32),
                                                               want.view(np.uint32)))
        if ref is not None:
            par["vs_reference_std_sort"] = max_rel(image, ref)
            par["vs_reference_std_sort_abs_over_max"] = max_abs_over_max(image, ref)
            par["pixels_differing_from_reference"] = int(
                (image.view(np.uint32) != ref.view(np.uint32)).any(axis=2).sum())
        par["z_ties"] = z_ties(p, p.type == 0)
        f64, _ = oracle_images(O, field, view, [pal, pal], [0, 1], [0.5, 2.0], True)
        par["vs_fp64"] = max_rel(image, f64)
        par["vs_fp64_abs_over_max"] = max_abs_over_max(image, f64)
        c1 = par
        par = {"c1": c1}
        if "c0" in out:
            par["c0"] = out["c0"]["parity"]
        if "fwd_bwd" in out:
            par["c3"] = {k: {"vs_fp64": v["parity_fwd_vs_fp64"],
                             "vs_fp64_abs_over_max": v["parity_fwd_vs_fp64_abs_over_max"]}
                         for k, v in out["fwd_bwd"].items()}
        out["parity"] = par
        log("parity %s" % par)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

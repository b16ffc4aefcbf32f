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
                frames); traffic from rocprofv3 FETCH_SIZE + WRITE_SIZE when
                --traffic-kib gives them (KiB per dispatch, already scaled);
  valu          the limit the splat actually meets: VALU instructions of the
                ordered splat per frame (rocprofv3 SQ_INSTS_VALU, committed
                summary) over the ordered splat stage's live duration, against
                the issue peak (256 CUs x 4 SIMDs x one wave64 instruction per
                4 clocks at 2.4 GHz);
  batched       the batched workload (64 views, synth.batch_jobs) as one
                splhip_render_batched call vs one splhip_render per view;
  fwd_bwd       the differentiable workload: 2M particles, 512^2, additive,
                and a 500k ordered variant; loss sum (pic - target)^2 against
                the render of a perturbed copy; forward vs forward + backward;
  epilogue      the post-processing of the frame (-expm1, 8-bit quantization,
                TGA byte order; splhip_postprocess) as one kernel: algorithmic
                bytes 12 B read + 3 B written per pixel over its event time,
                and the host copy of the 8-bit frame vs the float frame;
  normalize     particle_normalize of the frame's 16M particles (two passes:
                transform + per-type ranges, then clamp), bytes 13 B read in
                the first and 9 B read + 8 B written in the second per
                particle, over its event time;
  ingest        the frame's 16M particles written as a block binary file
                (bin_reader_block layout: six big-endian float columns) and
                read into device SoA arrays by splhip_read_bin_block (file in
                the page cache): bytes of the file per second;
  fragments     the frame's per-pixel updates (render_new's fragments, counted
                by the oracle: `bench.py --count-fragments`, committed summary)
                over the frame time and over the ordered splat stage;
  hbm_copy      a measured device copy (1 GiB) beside the 8 TB/s datasheet peak;
  cpu_baseline  the genuine reference host renderer (oracle/_ref, built from
                the Splotch sources) — or the oracle restatement if that
                library is absent — timed on this host's cores on a bounded
                random subsample of the same frame (same image, same params).
"""
from __future__ import annotations

import argparse
import contextlib
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/frame & Mparticles/s (16M, 2048²); batched and fwd+bwd vs plain"
HBM_PEAK_GBS = 8000.0
VALU_PEAK = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions per second (one per SIMD per 4 clocks)


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


def cpu_baseline(field, view, pal, budget_s=20.0):
    """Reference host renderer on a random subsample; returns the JSON object."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    from splotch_amd import synth

    use_ref = O.ref_available()
    n = field.n
    rng = np.random.default_rng(7)

    def sample(frac):
        if frac >= 1.0:
            return field
        m = max(2, int(n * frac))
        idx = np.sort(rng.choice(n, size=m, replace=False))
        f = synth.Field(field.x[idx], field.y[idx], field.z[idx], field.hsml[idx],
                        field.inten[idx], field.col[idx], field.type[idx])
        return f

    def run(f):
        t0 = time.perf_counter()
        for ptype, aeq, br in ((0, False, 0.5), (1, True, 2.0)):
            sel = f.type == ptype
            g = synth.Field(f.x[sel], f.y[sel], f.z[sel], f.hsml[sel], f.inten[sel], f.col[sel],
                            np.zeros(int(sel.sum()), np.uint8))
            if use_ref:
                params = {"xres": view.xres, "yres": view.yres, "fov": view.fov,
                          "a_eq_e": aeq, "brightness0": br}
                with quiet_stdout():
                    O.ref_host_rendering(g, view, [pal], params)
            else:
                O.host_rendering(g, view, [pal], a_eq_e=aeq, brightness=[br], stable=False)
        return time.perf_counter() - t0

    frac0 = 1.0 / 32
    t0 = run(sample(frac0))
    frac = min(1.0, frac0 * budget_s / max(t0, 1e-3))
    if frac > 0.6:  # the whole frame fits the budget: time exactly the GPU workload
        frac = 1.0
    f = sample(frac)
    t = run(f)
    threads = int(O.lib().orc_num_threads())
    return {"value": round(f.n / t / 1e6, 4), "unit": "Mparticles/s", "cores": threads,
            "kind": "reference" if use_ref else "port",
            "sample": "random %.4f subsample (%d of %d particles, both types), same 2048x2048 "
                      "frame and params, one run after a 1/32 calibration run, %.1f s; "
                      "%s host renderer with OpenMP on %d threads (%s)"
                      % (frac, f.n, n, t, "genuine Splotch" if use_ref else "oracle restatement",
                         threads, cpu_model()),
            "seconds": round(t, 3)}


def bench_batched(S, R, torch, steps, warmup, n_jobs=64):
    """Batched workload: n_jobs views (sizes, cameras and modes of
    synth.batch_jobs) rendered as one batch, and the same views one
    splhip_render call after another.  Aggregate Mparticles/s of each."""
    from splotch_amd import synth
    jobs = synth.batch_jobs(n_jobs)
    views = []
    for bj in jobs:
        dev = S.to_device(bj.field)
        cam = S.camera_setup(bj.view)
        prm = S.params(a_eq_e=bj.additive)
        pic = torch.empty(bj.view.xres * bj.view.yres * 3, dtype=torch.float32, device="cuda")
        views.append((dev, cam, prm, pic))
    ntot = sum(bj.field.n for bj in jobs)

    def plain():
        for dev, cam, prm, pic in views:
            R.render(dev, cam, prm, pic)

    def batched():
        R.render_batched(views)

    tuned_views = [(d, c, S.params(a_eq_e=bj.additive, hip_autotune=1), p)
                   for (d, c, _, p), bj in zip(views, jobs)]

    def batched_tuned():
        R.render_batched(tuned_views)

    res = {}
    for name, fn in (("plain", plain), ("batched", batched), ("batched_autotuned", batched_tuned)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        res[name] = (ms, ntot / (ms * 1e-3) / 1e6)
        if name == "batched_autotuned":
            chosen = R.tuning()
    log("batched %d views, %d particles: plain %.2f ms, batched %.2f ms, autotuned %.2f ms (%s)" %
        (n_jobs, ntot, res["plain"][0], res["batched"][0], res["batched_autotuned"][0], chosen))
    return {"views": n_jobs, "particles": ntot,
            "sides": sorted({bj.view.xres for bj in jobs}),
            "ms_per_batch": round(res["batched"][0], 3),
            "mparticles_s": round(res["batched"][1], 3),
            "plain_ms_per_batch": round(res["plain"][0], 3),
            "plain_mparticles_s": round(res["plain"][1], 3),
            "batched_vs_plain": round(res["batched"][1] / res["plain"][1], 3),
            "autotuned_ms_per_batch": round(res["batched_autotuned"][0], 3),
            "autotuned_vs_plain": round(res["batched_autotuned"][1] / res["plain"][1], 3),
            "autotuned_choice": chosen}


def bench_backward(S, R, torch, steps, warmup):
    """Differentiable workload: forward, and forward + backward of
    L = sum (pic - target)^2, target = render of the same field with
    x += N(0, 0.05) per axis and I *= U(0.8, 1.2)."""
    import numpy as np
    from splotch_amd import synth
    out = {}
    for name, n, seed, aeq in (("additive_2M", 2_000_000, 1300, True),
                               ("ordered_500k", 500_000, 1301, False)):
        fld = synth.plummer_field(n, seed)
        rng = np.random.default_rng(1302)
        tf = synth.Field(*(getattr(fld, k) + rng.normal(0, 0.05, n).astype(np.float32)
                           for k in ("x", "y", "z")), fld.hsml,
                         (fld.inten * rng.uniform(0.8, 1.2, n)).astype(np.float32), fld.col)
        view = synth.default_view(512)
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

        res = {}
        for key, fn in (("fwd", fwd), ("fwd_bwd", fwd_bwd)):
            for _ in range(warmup):
                fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                fn()
            torch.cuda.synchronize()
            res[key] = (time.perf_counter() - t0) * 1e3 / steps
        log("fwd+bwd %s: fwd %.2f ms, fwd+bwd %.2f ms" % (name, res["fwd"], res["fwd_bwd"]))
        out[name] = {"particles": n, "image": [512, 512], "additive": aeq,
                     "ms_fwd": round(res["fwd"], 3), "ms_fwd_bwd": round(res["fwd_bwd"], 3),
                     "fwd_bwd_over_fwd": round(res["fwd_bwd"] / res["fwd"], 3)}
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
    """200M particles (synth.large_share), 4096^2, ordered gas pass (halo half)
    + additive pass (background half): plain frames, then 16 frames along an
    orbit (azimuth k * 22.5 deg, elevation 0, radius 150, sky +z) rendered one
    after another.  CPU baseline: the reference host renderer on a random 1/100
    subsample, extrapolated by particle count (labelled)."""
    import numpy as np
    from splotch_amd import synth
    t0 = time.time()
    field = synth.large_share()
    log("generated %d particles in %.1f s" % (field.n, time.time() - t0))
    view = synth.default_view(4096)
    pal = synth.default_palette()
    dev = S.to_device(field, "cuda")
    R = S.Renderer(0)
    R.set_palette(0, *pal)
    R.set_palette(1, *pal)
    prm = S.params(ptypes=2, a_eq_e=[0, 1], brightness=[0.5, 2.0])
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
        orbit.append(S.camera_setup(v))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in orbit:
        R.render(dev, c, prm, pic)
    torch.cuda.synchronize()
    ms_orbit = (time.perf_counter() - t0) * 1e3 / 16
    out = {"metric": METRIC, "value": round(field.n / (ms * 1e-3) / 1e6, 3),
           "unit": "Mparticles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "large_share_200M_4096sq", "particles": field.n,
                      "gas_ordered": int((field.type == 0).sum()),
                      "additive": int((field.type == 1).sum()), "image": [4096, 4096],
                      "tile_size": 50, "parallelism": "single GPU"},
           "stages_ms": {k: round(v, 3) for k, v in stages.items()}, "counts": counts,
           "orbit16_ms_per_frame": round(ms_orbit, 3),
           "orbit16_mparticles_s": round(field.n / (ms_orbit * 1e-3) / 1e6, 3)}
    R.close()
    if not args.no_cpu:
        del dev
        torch.cuda.empty_cache()
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        rng = np.random.default_rng(7)
        idx = np.sort(rng.choice(field.n, size=field.n // 100, replace=False))
        sub = synth.Field(field.x[idx], field.y[idx], field.z[idx], field.hsml[idx],
                          field.inten[idx], field.col[idx], field.type[idx])
        use_ref = O.ref_available()
        t0 = time.perf_counter()
        for ptype, aeq, br in ((0, False, 0.5), (1, True, 2.0)):
            sel = sub.type == ptype
            g = synth.Field(sub.x[sel], sub.y[sel], sub.z[sel], sub.hsml[sel], sub.inten[sel],
                            sub.col[sel], np.zeros(int(sel.sum()), np.uint8))
            if use_ref:
                with quiet_stdout():
                    O.ref_host_rendering(g, view, [pal], {"xres": 4096, "yres": 4096,
                                                          "fov": view.fov, "a_eq_e": aeq,
                                                          "brightness0": br})
            else:
                O.host_rendering(g, view, [pal], a_eq_e=aeq, brightness=[br], stable=False)
        t = time.perf_counter() - t0
        threads = int(O.lib().orc_num_threads())
        cpu_mps = sub.n / t / 1e6
        out["cpu_baseline"] = {
            "value": round(cpu_mps, 4), "unit": "Mparticles/s", "cores": threads,
            "kind": "reference" if use_ref else "port",
            "sample": "EXTRAPOLATED: random 1/100 subsample (%d particles, both types) of the "
                      "200M field rendered at 4096^2, %.1f s, rate taken as the full frame's"
                      % (sub.n, t),
            "speedup": round(out["value"] / cpu_mps, 2)}
    print(json.dumps(out), flush=True)


def count_fragments(args):
    """Fragments of the default frame per pass, from the oracle (CPU)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
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


def hbm_copy(torch, nbytes=1 << 30, reps=10):
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return {"bytes_moved": 2 * nbytes, "ms": round(ms, 4),
            "GBs": round(2 * nbytes / (ms * 1e-3) / 1e9, 1), "datasheet_GBs": HBM_PEAK_GBS}


def copy_calibration(torch, steps=20):
    """A plain device copy of 1 GiB, for scaling rocprofv3 FETCH_SIZE /
    WRITE_SIZE: each dispatch reads and writes exactly 1 GiB."""
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(steps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": 4 * n, "dispatches": steps}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-gas", type=int, default=8_000_000)
    ap.add_argument("--n-stars", type=int, default=8_000_000)
    ap.add_argument("--res", type=int, default=2048)
    ap.add_argument("--region", type=int, default=0, help="hip_region (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched workload")
    ap.add_argument("--no-backward", action="store_true", help="skip the fwd+bwd workload")
    ap.add_argument("--workload", default="gas_stars",
                    choices=["gas_stars", "large_share"],
                    help="gas_stars: 16M at 2048^2 (default); large_share: 200M at 4096^2")
    ap.add_argument("--count-fragments", action="store_true",
                    help="print the default frame's fragments per pass (oracle, CPU) and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (profiler counter calibration)")
    ap.add_argument("--traffic-kib", type=float, nargs=2, default=None,
                    metavar=("FETCH", "WRITE"),
                    help="scaled FETCH_SIZE and WRITE_SIZE (KiB) of one ordered k_splat8 "
                         "dispatch, from rocprofv3, reported as roofline.traffic")
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

    import numpy as np
    import torch

    if args.copy_calibration:
        copy_calibration(torch)
        return
    if args.workload == "large_share":
        import splotch_amd as S
        bench_large_share(S, torch, args)
        return

    import splotch_amd as S
    from splotch_amd import synth

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

    W, H = view.xres, view.yres
    splat_ms = stages["ordered_k_splat8"]
    splat_bytes = 24 * counts["ordered_active"] + 12 * W * H
    achieved = splat_bytes / (splat_ms * 1e-3) / 1e9
    traffic, traffic_src = None, None
    if args.traffic_kib:
        traffic = int(round((args.traffic_kib[0] + args.traffic_kib[1]) * 1024))
        traffic_src = "--traffic-kib"
    else:  # the committed rocprofv3 summary of this workload, if present
        prof = os.path.join(ROOT, "profiles", "pmc_traffic_gas_stars_16M_2048.md")
        default = (args.n_gas, args.n_stars, args.res) == (8_000_000, 8_000_000, 2048)
        if default and os.path.exists(prof):
            for line in open(prof):
                if line.startswith("traffic_bytes_per_dispatch k_splat8<0, 4>:"):
                    traffic = int(line.split(":")[1])
                    traffic_src = "profiles/pmc_traffic_gas_stars_16M_2048.md"
    valu = None
    vprof = os.path.join(ROOT, "profiles", "pmc_valu_gas_stars_16M_2048.md")
    if (args.n_gas, args.n_stars, args.res) == (8_000_000, 8_000_000, 2048) and \
            os.path.exists(vprof):
        for line in open(vprof):
            if line.startswith("valu_instructions_per_frame ordered_splat:"):
                ninst = int(line.split(":")[1])
                rate = ninst / (stages["ordered_splat"] * 1e-3)
                valu = {"kernels": "k_splat8<ordered> + k_splat_pc<ordered> (concurrent)",
                        "instructions": ninst, "stage_ms": stages["ordered_splat"],
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
                   "hip_region": prm.hip_region or 8, "parallelism": "single GPU"},
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
    fprof = os.path.join(ROOT, "profiles", "fragments_gas_stars_16M_2048.md")
    if (args.n_gas, args.n_stars, args.res) == (8_000_000, 8_000_000, 2048) and \
            os.path.exists(fprof):
        fr = {}
        for line in open(fprof):
            if line.startswith("fragments_"):
                k, v = line.split(":")
                fr[k.strip()] = int(v)
        tot = fr["fragments_ordered"] + fr["fragments_additive"]
        out["fragments"] = {
            "ordered": fr["fragments_ordered"], "additive": fr["fragments_additive"],
            "per_s_frame": round(tot / (ms_frame * 1e-3), 1),
            "ordered_per_s_splat_stage": round(fr["fragments_ordered"] /
                                               (stages["ordered_splat"] * 1e-3), 1),
            "source": "profiles/fragments_gas_stars_16M_2048.md"}
    out["hbm_copy"] = hbm_copy(torch)
    out["epilogue"] = bench_epilogue(S, R, torch, pic, W, H)
    out["normalize"] = bench_normalize(S, R, torch, dev)
    out["ingest"] = bench_ingest(S, R, torch, field)
    if not args.no_batched:
        out["batched"] = bench_batched(S, R, torch, max(2, args.steps // 2), 1)
    if not args.no_backward:
        out["fwd_bwd"] = bench_backward(S, R, torch, max(3, args.steps // 2), 1)
    if not args.no_cpu:
        del dev
        torch.cuda.empty_cache()
        cb = cpu_baseline(field, view, pal, args.cpu_budget)
        cb["speedup"] = round(out["value"] / cb["value"], 2)
        out["cpu_baseline"] = cb
    R.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

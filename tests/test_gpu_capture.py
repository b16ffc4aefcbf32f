"""Batched renders with device-side counts, captured into a HIP graph.

splhip_capture_reserve renders a batch once with host waits, sizes the pass
capacities from it and renders it once more with device-side counts (no host
wait, allocation, host copy or memset).  Under stream capture the same
splhip_render_batched call queues that pipeline, and the graph is replayed.
Bar: every image of the reserve and of every replay equals the plain render
of its view bit for bit (also with staged long lists, whose partition against
producer/consumer workgroups and the 8x8 splat is then made on the device); replays after the particle values change (same
arrays, same counts) equal the plain renders of the new values; a replay
whose entries exceed the reserved capacity reports it (splhip_capture_status)
and writes nothing past its buffers.  Replays are interleaved with eager work
and synchronisations, and each replay is checked.
"""
import math
import sys

import numpy as np
import pytest

from helpers import assert_bitexact
from splotch_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def R():
    import splotch_amd as S
    assert torch.cuda.is_available(), "GPU test without a GPU"
    r = S.Renderer(0)
    pal = synth.default_palette()
    for t in range(2):
        r.set_palette(t, *pal)
    yield r
    torch.cuda.synchronize()
    r.close()


def make_views(S, jobs, prm_kw=None):
    views = []
    for bj in jobs:
        views.append((S.to_device(bj.field), S.camera_setup(bj.view),
                      S.params(a_eq_e=bj.additive, **(prm_kw or {})),
                      torch.empty(bj.view.xres * bj.view.yres * 3, dtype=torch.float32,
                                  device="cuda")))
    return views


def plain_images(R, views):
    out = []
    for dev, cam, prm, pic in views:
        p = torch.empty_like(pic)
        R.render(dev, cam, prm, p)
        out.append(p)
    torch.cuda.synchronize()
    return [p.cpu().numpy() for p in out]


def check(views, want, what):
    torch.cuda.synchronize()
    for j, (_, _, _, pic) in enumerate(views):
        assert_bitexact(pic.cpu().numpy(), want[j], "%s view %d" % (what, j))


def new_values(views, jobs, seed):
    """Overwrite every view's particle arrays in place with another field of
    the same count (positions, sizes, intensities and colours change)."""
    for (dev, _, _, _), bj in zip(views, jobs):
        f = synth.plummer_field(bj.field.n, seed + bj.field.n % 1000)
        for k in ("x", "y", "z", "hsml", "inten", "col"):
            dev[k].copy_(torch.from_numpy(getattr(f, k)))


@pytest.mark.parametrize("prm_kw", [dict(), dict(hip_heavy=1, hip_heavy_max=64),
                                    dict(hip_stage=64, hip_stage_ordered=64),
                                    dict(hip_stage=256, hip_stage_ordered=128, hip_heavy=64,
                                         hip_heavy_max=4)])
def test_capture_replay_bitexact(R, prm_kw):
    import splotch_amd as S
    log("case capture: 12 views of batch_jobs(seed 1200, 1/20 scale)", prm_kw)
    jobs = synth.batch_jobs(12, n_scale=0.05)
    views = make_views(S, jobs, prm_kw)
    want = plain_images(R, views)
    R.capture_reserve(views, 1.5)
    check(views, want, "reserve (device counts)")
    assert R.capture_status() == 0
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            R.render_batched(views)
    torch.cuda.synchronize()
    log("capture: graph captured")
    for rep in range(3):
        for v in views:  # eager work between replays
            v[3].fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        check(views, want, "replay %d" % rep)
        assert R.capture_status() == 0
        plain_images(R, views[:2])  # eager renders on the same context in between
    # new particle values, same arrays and counts
    new_values(views, jobs, 2400)
    want2 = plain_images(R, views)
    g.replay()
    check(views, want2, "replay with new particle values")
    st = R.capture_status()
    log("capture: new values, overflow bits %d" % st)
    assert st == 0
    del g


def test_capture_overflow_flag(R):
    """Particles grown threefold after the reservation (headroom 1): the bin
    entries exceed the reserved capacity; the replay reports it and leaves
    the passes out, and a new reservation renders the new values exactly."""
    import splotch_amd as S
    log("case capture overflow")
    jobs = synth.batch_jobs(4, seed=1250, n_scale=0.05)
    views = make_views(S, jobs)
    R.capture_reserve(views, 1.0)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            R.render_batched(views)
    torch.cuda.synchronize()
    for dev, _, _, _ in views:
        dev["hsml"].mul_(3.0)
    torch.cuda.synchronize()
    g.replay()
    st = R.capture_status()
    log("capture overflow: bits %d" % st)
    assert st != 0
    want = plain_images(R, views)
    R.capture_reserve(views, 1.25)
    check(views, want, "re-reserved")
    del g


def test_capture_errors(R):
    import splotch_amd as S
    jobs = synth.batch_jobs(2, seed=1260, n_scale=0.02)
    views = make_views(S, jobs)
    R.capture_reserve(views, 1.25)
    other = make_views(S, synth.batch_jobs(2, seed=1270, n_scale=0.02))
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with pytest.raises(Exception):   # a batch other than the reserved one
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                R.render_batched(other)
    torch.cuda.synchronize()
    with pytest.raises(S.SplhipError):   # 16 px work regions
        R.capture_reserve(make_views(S, jobs, dict(hip_region=16)), 1.25)

"""GPU parity of the batched render (splhip_render_batched), through the C-ABI.

Bar: every view of a batch is bit-identical to splhip_render of that view
alone, and (on small views) to the oracle's host_rendering with the stable
(z, index) depth order.  Batches mix additive and ordered views, image sizes,
tile sizes, cameras, empty and fully culled views, and views with both
particle kinds.
"""
import math
import sys

import numpy as np
import pytest

import pyoracle as O
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


def _views(jobs, **prm_kw):
    import splotch_amd as S
    out = []
    for j, bj in enumerate(jobs):
        dev = S.to_device(bj.field)
        cam = S.camera_setup(bj.view)
        types = bj.field.type
        if types is not None and (types == 1).any():
            prm = S.params(ptypes=2, a_eq_e=[0, 1], brightness=[0.5, 2.0], **prm_kw)
        else:
            prm = S.params(a_eq_e=bj.additive, **prm_kw)
        pic = torch.full((bj.view.xres * bj.view.yres * 3,), float("nan"), dtype=torch.float32,
                         device="cuda")
        out.append((dev, cam, prm, pic))
    return out


def _plain(R, views):
    imgs = []
    for dev, cam, prm, pic in views:
        p = torch.full_like(pic, float("nan"))
        R.render(dev, cam, prm, p)
        imgs.append(p)
    torch.cuda.synchronize()
    return [p.cpu().numpy() for p in imgs]


def _small_batch(seed=2200, n_jobs=12):
    jobs = synth.batch_jobs(n_jobs, seed=seed, n_min=2e3, n_max=6e4)
    # shrink the images so the oracle stays fast
    for bj in jobs:
        bj.view.xres = bj.view.yres = bj.view.xres // 4
    return jobs


def test_batched_equals_plain(R):
    jobs = _small_batch()
    views = _views(jobs)
    log("batched vs plain:", [(bj.field.n, bj.view.xres, bj.additive) for bj in jobs])
    R.render_batched(views)
    torch.cuda.synchronize()
    log("region entries of the batch's passes:", R.binning())
    got = [v[3].cpu().numpy() for v in views]
    ref = _plain(R, views)
    for j, (g, r) in enumerate(zip(got, ref)):
        assert_bitexact(g, r, "job %d batched vs plain" % j)


def test_batched_vs_oracle(R):
    jobs = _small_batch(seed=2300, n_jobs=4)
    views = _views(jobs)
    R.render_batched(views)
    torch.cuda.synchronize()
    pal = synth.default_palette()
    for j, (bj, v) in enumerate(zip(jobs, views)):
        opic, _ = O.host_rendering(bj.field, bj.view, [pal], a_eq_e=bj.additive, stable=True)
        got = v[3].cpu().numpy().reshape(bj.view.xres, bj.view.yres, 3)
        assert_bitexact(got, opic, "job %d vs oracle" % j)


@pytest.mark.parametrize("region", [0, 16, 64])
def test_batched_mixed_edge_views(R, region):
    """Tile sizes, both particle kinds in one view, an empty view, a view whose
    particles are all behind the camera, and a view repeated twice, with 8 px
    work regions and larger ones (the region splat looks up each unit's view)."""
    import splotch_amd as S
    log("case edge views, hip_region", region)
    base = _small_batch(seed=2400, n_jobs=3)
    gas = synth.plummer_field(8000, 2401, ptype=0)
    stars = synth.plummer_field(8000, 2402, hsml_fixed=0.05, ptype=1)
    mixed = synth.BatchJob(synth.concat(gas, stars), synth.orbit_view(0.3, 0.2, 150.0, 160), False)
    empty = synth.BatchJob(synth.plummer_field(1, 2403), synth.orbit_view(1.0, 0.1, 150.0, 96),
                           True)
    empty.field = synth.Field(*[np.zeros(0, np.float32)] * 6)
    behind = synth.BatchJob(synth.plummer_field(5000, 2404),
                            synth.View((50.0, 50.0, 300.0), (50.0, 50.0, 400.0), (0.0, 1.0, 0.0),
                                       128, 100, 30.0), False)
    jobs = base + [mixed, empty, behind, base[1]]
    views = _views(jobs, hip_region=region)
    # per-view tile sizes (a render key of each view)
    for (dev, cam, prm, pic), T in zip(views, [50, 16, 37, 64, 50, 8, 50]):
        prm.tile_size = T
    log("edge batch sizes", [bj.field.n for bj in jobs])
    R.render_batched(views)
    torch.cuda.synchronize()
    got = [v[3].cpu().numpy() for v in views]
    ref = _plain(R, views)
    for j, (g, r) in enumerate(zip(got, ref)):
        assert_bitexact(g, r, "edge job %d batched vs plain" % j)
    assert not np.any(got[4]), "empty view must be black"
    assert not np.any(got[5]), "fully culled view must be black"
    assert_bitexact(got[1], got[6], "repeated view")


def test_batched_region_binning(R):
    """Views of large-footprint particles bin through regions of 8x8 blocks, numbered
    per view from the view's first region: every image equals its plain
    render and, for the first view, the oracle."""
    tiles, res = [50, 37, 24, 50], [160, 144, 176, 120]
    jobs = []
    for k in range(4):
        fld = synth.plummer_field(12_000, 2500 + k, halo_frac=1.0, n_halos=2,
                                  hsml_mu=math.log(3.0))
        jobs.append(synth.BatchJob(fld, synth.orbit_view(0.4 * k, 0.2, 150.0, res[k]), k % 2 == 1))
    views = _views(jobs)
    for (dev, cam, prm, pic), T in zip(views, tiles):
        prm.tile_size = T
    R.render_batched(views)
    torch.cuda.synchronize()
    regions = R.binning()
    log("region entries of the batch's passes:", regions)
    assert regions[0] > 0 and regions[1] > 0
    got = [v[3].cpu().numpy() for v in views]
    ref = _plain(R, views)
    for j, (g, r) in enumerate(zip(got, ref)):
        assert_bitexact(g, r, "region job %d batched vs plain" % j)
    pal = synth.default_palette()
    opic, _ = O.host_rendering(jobs[0].field, jobs[0].view, [pal], a_eq_e=False, stable=True)
    assert_bitexact(got[0].reshape(opic.shape), opic, "region job 0 vs oracle")


@pytest.mark.parametrize("variant,split,heavy,hmax,ppw,region",
                         [(4, 0, 0, 0, 0, 0), (1, 0, -1, 0, 0, 0), (3, 64, 200, 0, 0, 0),
                          (4, 64, 1, 0, 0, 0), (4, 0, 300, 1 << 30, 0, 0), (4, 64, -1, 0, 1, 0),
                          (4, 0, -1, 0, 2, 0), (4, 64, 1, 0, 4, 0), (4, 0, -1, 0, 8, 0),
                          (4, 0, 0, 0, 0, 16), (4, 0, 0, 0, 0, 32), (4, 0, 0, 0, 0, 64)])
def test_batched_tuning_invariant(R, variant, split, heavy, hmax, ppw, region):
    """The tuning keys (the 8x8 splat's decomposition, particles per wave,
    and the work region of 8-64 px) change the GPU work decomposition only."""
    log("case batched tuning", variant, split, heavy, hmax, ppw, region)
    jobs = _small_batch(seed=2500, n_jobs=6)
    views0 = _views(jobs)
    R.render_batched(views0)
    views1 = _views(jobs, hip_variant=variant, hip_split=split, hip_heavy=heavy,
                    hip_heavy_max=hmax, hip_ppw=ppw, hip_region=region)
    R.render_batched(views1)
    torch.cuda.synchronize()
    for j, (a, b) in enumerate(zip(views0, views1)):
        assert_bitexact(b[3].cpu().numpy(), a[3].cpu().numpy(), "job %d tuned" % j)


def test_batched_errors(R):
    import splotch_amd as S
    jobs = _small_batch(seed=2600, n_jobs=2)
    views = _views(jobs, hip_region=24)
    with pytest.raises(S.SplhipError):
        R.render_batched(views)
    with pytest.raises(S.SplhipError):
        R.render_batched(_views(jobs, hip_ppw=3))
    R.render_batched([])  # no views: nothing to do


def test_batched_autotune(R):
    """hip_autotune times the work regions (8-64 px) and the 8x8 splat's
    decomposition (producer/consumer cap, channel split, inner loop and
    particles per wave) on the first render of a batch bucket, then reuses
    the choice, also for another batch of the same bucket; images are those
    of the default."""
    jobs = _small_batch(seed=2700, n_jobs=6)
    ref = _views(jobs)
    R.render_batched(ref)
    first = R.tuning()
    assert not first["autotuned"]
    chosen = None
    for call in range(2):
        views = _views(jobs, hip_autotune=1)
        R.render_batched(views)
        t = R.tuning()
        log("autotune call %d chose" % call, t)
        assert t["autotuned"]
        assert t["region"] in (8, 16, 32, 64)
        if t["region"] == 8:
            assert t["variant"] in (3, 4) and t["split"] in (4096, 8192, 32768)
            assert t["heavy_max"] in (0, 96, 192, 256) and t["ppw"] in (0, 1, 2, 4, 8)
        assert chosen is None or t == chosen
        chosen = t
        torch.cuda.synchronize()
        for j, (a, b) in enumerate(zip(views, ref)):
            assert_bitexact(a[3].cpu().numpy(), b[3].cpu().numpy(), "autotuned job %d" % j)


def test_scene_file_frames(R):
    """The frames of a scene file (12 cameras on an orbit, particles changing
    every third frame by fidx) rendered group by group through
    splhip_render_batched: each image equals the plain render of its camera."""
    import splotch_amd as S
    rows = []
    for k in range(12):
        a = 0.5 * k
        rows.append("%.5f %.5f %.5f 50 50 50 0 0 1 %d" % (50 + 150 * math.cos(a),
                                                         50 + 150 * math.sin(a), 60.0, k // 3))
    text = "\n".join(rows) + "\n"
    par = "xres=160\nyres=128\nfov=30\na_eq_e=false\n"
    fields = {f: S.to_device(synth.plummer_field(20_000, 2800 + f, hsml_mu=math.log(0.5)))
              for f in range(4)}
    imgs = R.render_scene_file(lambda fidx: fields[int(round(fidx))], par, text)
    torch.cuda.synchronize()
    frames = S.scene_views(par, text)
    assert len(imgs) == 12 and [f[3] for f in frames][:3] == [True, True, False]
    for k, ((cam, prm, fidx, _), img) in enumerate(zip(frames, imgs)):
        p = torch.empty_like(img)
        R.render(fields[int(round(fidx))], cam, prm, p)
        torch.cuda.synchronize()
        assert_bitexact(img.cpu().numpy(), p.cpu().numpy(), "scene frame %d" % k)

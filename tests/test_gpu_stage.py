"""GPU parity of the staged long lists, through the C-ABI.

A list from hip_stage (additive pass) or hip_stage_ordered (ordered pass)
entries on has its per-pixel operands computed across the GPU first
(k_stage_ops), then applied in list order by one wave per list, channel and
16-pixel slice (k_stage_chain).  Bar: bit-exact against the genuine fixtures
and the oracle, and against the unstaged render, with every list staged,
with staged lists beside producer/consumer workgroups and normal items, in
chunked passes (chains continued from the image), for both passes of one
view, in batches of views, with non-finite particle intensities (NaN in
exactly the oracle's pixels), with a non-finite q_c (gray_absorption 0:
no ordered staging, and the image unchanged), and with long lists that
neither the staged nor the producer/consumer kernels take, walked by channel
(plain and replayed from a captured device-count render).
"""
import math
import sys

import numpy as np
import pytest

import pyoracle as O
from helpers import assert_bitexact, field_from, view_from
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
    yield r
    torch.cuda.synchronize()
    r.close()


def render(R, field, view, pals, prm):
    import splotch_amd as S
    for t, (x, rgb) in enumerate(pals):
        R.set_palette(t, x, rgb)
    cam = S.camera_setup(view)
    dev = S.to_device(field)
    pic = torch.empty(view.xres * view.yres * 3, dtype=torch.float32, device="cuda")
    R.render(dev, cam, prm, pic)
    torch.cuda.synchronize()
    return pic.cpu().numpy().reshape(view.xres, view.yres, 3)


@pytest.mark.parametrize("T", [50, 16, 8, 64])
def test_staged_fixtures(R, golden, T):
    """Every non-empty block staged (hip_stage 1), against the genuine build's
    raw image, for each reference tile size (blocks cut at tile edges)."""
    import splotch_amd as S
    log("case staged fixture T", T)
    f = golden("additive_small")
    p = golden("palette")
    pic = render(R, field_from(f), view_from(f), [(p["x"], p["rgb"])],
                 S.params(a_eq_e=True, tile_size=T, hip_stage=1))
    assert_bitexact(pic, f["pic_T%d" % T], "staged fixture T=%d" % T)


@pytest.mark.parametrize("g05", [False, True])
def test_staged_ordered_fixture(R, golden, g05):
    """Every non-empty block of the ordered fixture staged, against the
    genuine build's raw image in the stable depth order (gray_absorption 0.2
    and 0.5)."""
    import splotch_amd as S
    log("case staged ordered fixture, gray 0.5:", g05)
    f = golden("ordered_small")
    p = golden("palette")
    kw = dict(gray_absorption=0.5) if g05 else {}
    pic = render(R, field_from(f), view_from(f), [(p["x"], p["rgb"])],
                 S.params(a_eq_e=False, brightness=[float(f["brightness"][0])],
                          hip_stage_ordered=1, **kw))
    assert_bitexact(pic, f["pic_stable_T50_g05" if g05 else "pic_stable_T50"],
                    "staged ordered fixture")


def test_staged_ordered_dense_cluster_vs_oracle(R):
    """Ordered: staged at several thresholds, beside producer/consumer
    workgroups and channel items, in 8 px and larger tiles."""
    import splotch_amd as S
    fld = synth.plummer_field(60_000, 113, halo_frac=1.0, n_halos=1, hsml_mu=math.log(0.8))
    v = synth.default_view(160, 144)
    pal = synth.default_palette()
    opic, _ = O.host_rendering(fld, v, [pal], a_eq_e=False, stable=True)
    for stage, heavy, hmax, split, T in ((1, 0, 0, 0, 50), (64, 0, 0, 0, 50),
                                         (2048, 64, 0, 0, 50), (512, 16, 3, 32, 37),
                                         (4096, -1, 0, 16, 16), (-1, 0, 0, 0, 50)):
        log("case ordered dense cluster staged", stage, heavy, hmax, split, T)
        pic = render(R, fld, v, [pal], S.params(a_eq_e=False, hip_stage_ordered=stage,
                                                hip_heavy=heavy, hip_heavy_max=hmax,
                                                hip_split=split, tile_size=T))
        if T == 50:
            assert_bitexact(pic, opic, "ordered dense cluster stage=%d heavy=%d max=%d split=%d"
                            % (stage, heavy, hmax, split))
        else:
            ref = render(R, fld, v, [pal], S.params(a_eq_e=False, hip_stage_ordered=-1,
                                                    tile_size=T))
            assert_bitexact(pic, ref, "ordered dense cluster T=%d stage=%d" % (T, stage))


def test_staged_ordered_gray_zero(R):
    """gray_absorption 0 gives q_c = 0/0 for channels with e_c = 0: no list is
    staged (it would spread the NaN over whole blocks), and the image equals
    the unstaged render, NaN in the same pixels."""
    import splotch_amd as S
    fld = synth.plummer_field(20_000, 1620, halo_frac=1.0, n_halos=3, hsml_mu=math.log(0.6))
    v = synth.default_view(176, 144)
    pal = synth.default_palette()
    a = render(R, fld, v, [pal], S.params(a_eq_e=False, gray_absorption=0.0, hip_stage_ordered=1))
    b = render(R, fld, v, [pal], S.params(a_eq_e=False, gray_absorption=0.0,
                                          hip_stage_ordered=-1))
    assert np.isnan(a).any()
    assert_bitexact(a, b, "gray_absorption 0, staging requested")


def test_staged_dense_cluster_vs_oracle(R):
    """A dense cluster (lists of thousands of particles per 8x8 block): staged
    at several thresholds, beside producer/consumer workgroups (hip_heavy
    below hip_stage) and per-channel items, and unstaged."""
    import splotch_amd as S
    fld = synth.plummer_field(60_000, 113, halo_frac=1.0, n_halos=1, hsml_mu=math.log(0.8))
    v = synth.default_view(160, 144)
    pal = synth.default_palette()
    opic, _ = O.host_rendering(fld, v, [pal], a_eq_e=True)
    for stage, heavy, hmax, split, var in ((1, 0, 0, 0, 0), (64, 0, 0, 0, 0),
                                           (256, 64, 0, 0, 0), (1024, 16, 3, 32, 0),
                                           (4096, -1, 0, 16, 3), (256, -1, 0, 0, 1),
                                           (-1, 0, 0, 0, 0), (0, 0, 0, 0, 0)):
        log("case dense cluster staged", stage, heavy, hmax, split, var)
        pic = render(R, fld, v, [pal], S.params(a_eq_e=True, hip_stage=stage, hip_heavy=heavy,
                                                hip_heavy_max=hmax, hip_split=split,
                                                hip_variant=var))
        assert_bitexact(pic, opic, "dense cluster stage=%d heavy=%d max=%d split=%d var=%d"
                        % (stage, heavy, hmax, split, var))


def test_staged_trace_and_profile(R):
    """The work-item trace has a record per staged list (all three channels,
    its length) and the additive heavy stage a time; the image is unchanged."""
    import splotch_amd as S
    fld = synth.plummer_field(60_000, 114, halo_frac=1.0, n_halos=1, hsml_mu=math.log(0.8))
    v = synth.default_view(160, 144)
    pal = synth.default_palette()
    prm = S.params(a_eq_e=True, hip_stage=512)
    want = render(R, fld, v, [pal], S.params(a_eq_e=True, hip_stage=-1))
    R.set_profiling(True)
    R.set_trace(True)
    try:
        got = render(R, fld, v, [pal], prm)
        st, _ = R.stage_times()
        tr = R.trace_items(False)
    finally:
        R.set_trace(False)
        R.set_profiling(False)
    assert_bitexact(got, want, "traced staged frame")
    staged = tr[(tr[:, 2] & 0xffffffff) >= 512]
    log("staged lists traced: %d, stages %s" % (len(staged), st))
    assert len(staged) > 0
    assert np.all((staged[:, 2] >> 32) == 3)
    assert np.all(tr[:, 1] >= tr[:, 0])
    assert st["additive_staged"] > 0


@pytest.mark.parametrize("chunk", [20_000, 3_000])
def test_staged_chunked_passes(R, chunk):
    """Chunked passes continue each staged chain from the image (and, for the
    additive pass on top of an ordered one, from a scratch image): equal to
    the unchunked, unstaged render."""
    import splotch_amd as S
    log("case staged chunked", chunk)
    gas = synth.plummer_field(30_000, 1500, ptype=0, hsml_mu=math.log(0.8))
    stars = synth.plummer_field(30_000, 1501, ptype=1, hsml_mu=math.log(0.8))
    fld = synth.concat(gas, stars)
    v = synth.default_view(192, 160)
    pal = synth.default_palette()
    for aeq in ([0, 1], [1, 1], [2, 1]):
        want = render(R, fld, v, [pal, pal], S.params(ptypes=2, a_eq_e=aeq,
                                                      brightness=[0.5, 2.0], hip_stage=-1,
                                                      hip_stage_ordered=-1))
        for stage in (32, 300):
            got = render(R, fld, v, [pal, pal],
                         S.params(ptypes=2, a_eq_e=aeq, brightness=[0.5, 2.0], hip_stage=stage,
                                  hip_stage_ordered=stage, hip_chunk_entries=chunk))
            _, counts = R.stage_times()
            assert counts["additive_entries"] > chunk
            assert_bitexact(got, want, "staged chunked %s stage %d" % (aeq, stage))
            got = render(R, fld, v, [pal, pal],
                         S.params(ptypes=2, a_eq_e=aeq, brightness=[0.5, 2.0], hip_stage=stage,
                                  hip_stage_ordered=stage))
            assert_bitexact(got, want, "staged %s stage %d" % (aeq, stage))


def test_staged_batched_views(R):
    """A batch of additive and ordered views of several image sizes and tile
    sizes: with staged lists every image equals its unstaged plain render."""
    import splotch_amd as S
    pal = synth.default_palette()
    for t in range(2):
        R.set_palette(t, *pal)
    jobs = synth.batch_jobs(8, seed=2900, n_min=2e4, n_max=8e4)
    for bj in jobs:
        bj.view.xres = bj.view.yres = bj.view.xres // 4
    tiles = [50, 37, 16, 64, 50, 24, 50, 8]
    views, plain = [], []
    for bj, T in zip(jobs, tiles):
        dev = S.to_device(bj.field)
        cam = S.camera_setup(bj.view)
        n = bj.view.xres * bj.view.yres * 3
        p = torch.empty(n, dtype=torch.float32, device="cuda")
        R.render(dev, cam, S.params(a_eq_e=bj.additive, tile_size=T, hip_stage=-1,
                                    hip_stage_ordered=-1), p)
        plain.append(p)
        views.append((dev, cam, S.params(a_eq_e=bj.additive, tile_size=T, hip_stage=64,
                                         hip_stage_ordered=64),
                      torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")))
    log("staged batch:", [(bj.field.n, bj.view.xres, bj.additive) for bj in jobs])
    assert any(bj.additive for bj in jobs) and not all(bj.additive for bj in jobs)
    R.render_batched(views)
    torch.cuda.synchronize()
    for j, (v, p) in enumerate(zip(views, plain)):
        assert_bitexact(v[3].cpu().numpy(), p.cpu().numpy(), "staged batch view %d" % j)


def test_staged_non_finite(R):
    """Particles with infinite or NaN intensity: an infinite term meets a
    zero Gaussian factor inside the footprint (NaN) and nothing outside it;
    the staged sums have NaN in exactly the oracle's pixels and every other
    value bit for bit."""
    import splotch_amd as S
    fld = synth.plummer_field(20_000, 1630, halo_frac=1.0, n_halos=1, hsml_mu=math.log(0.8))
    rng = np.random.default_rng(1631)
    bad = rng.choice(fld.n, size=40, replace=False)
    fld.inten[bad[:20]] = np.inf
    fld.inten[bad[20:]] = np.nan
    v = synth.default_view(160, 144)
    pal = synth.default_palette()
    opic, _ = O.host_rendering(fld, v, [pal], a_eq_e=True)
    pic = render(R, fld, v, [pal], S.params(a_eq_e=True, hip_stage=16))
    on = np.isnan(opic)
    log("NaN channels: oracle %d of %d" % (int(on.sum()), on.size))
    assert on.any()
    assert np.array_equal(np.isnan(pic), on), "NaN pixels differ"
    assert_bitexact(pic[~on], opic[~on], "finite pixels, staged non-finite intensities")


def test_staged_deterministic(R):
    import splotch_amd as S
    fld = synth.plummer_field(150_000, 1640, halo_frac=0.9, n_halos=2)
    v = synth.default_view(256, 224)
    pal = synth.default_palette()
    a = render(R, fld, v, [pal], S.params(a_eq_e=True, hip_stage=128))
    b = render(R, fld, v, [pal], S.params(a_eq_e=True, hip_stage=128))
    c = render(R, fld, v, [pal], S.params(a_eq_e=True, hip_stage=-1))
    assert_bitexact(a, b, "staged repeat")
    assert_bitexact(a, c, "staged vs unstaged")


def test_leftover_long_lists_by_channel(R):
    """Long lists that neither the staged nor the producer/consumer kernels
    take are walked by channel, three waves each: 1M gas particles at 512^2
    stage their lists from 11584 entries on (the longest / 16, rounded up to a
    level), 4 producer/consumer workgroups take the next 4, and the other
    lists from 8192 entries on (the split level for 3.9M entries) are left.
    The trace holds three one-channel records for each of them; the image
    equals the oracle's bit for bit, plain and replayed from a captured
    device-count render."""
    import splotch_amd as S
    fld = synth.plummer_field(1_000_000, 1100, hsml_mu=math.log(0.5), ptype=0)
    v = synth.default_view(512)
    pal = synth.default_palette()
    log("case leftover long lists by channel: 1M gas, 512^2, hip_heavy_max 4")
    want, _ = O.host_rendering(fld, v, [pal], a_eq_e=False, stable=True, brightness=[0.5])
    prm = S.params(a_eq_e=False, brightness=[0.5], hip_heavy_max=4)
    R.set_trace(True)
    try:
        got = render(R, fld, v, [pal], prm)
        tr = R.trace_items(True)
    finally:
        R.set_trace(False)
    assert_bitexact(got, want, "leftover long lists by channel")
    ln = (tr[:, 2] & 0xffffffff).astype(np.int64)
    ch = tr[:, 2] >> 32
    split = (ch != 3) & (ln >= 8192)
    log("one-channel records of long lists: %d" % int(split.sum()))
    assert split.sum() >= 3 and split.sum() % 3 == 0
    for c in range(3):
        assert ((ch == c) & (ln >= 8192)).sum() == split.sum() // 3
    # the same view through a captured device-count render
    cam = S.camera_setup(v)
    views = [(S.to_device(fld), cam, prm,
              torch.empty(v.xres * v.yres * 3, dtype=torch.float32, device="cuda"))]
    R.capture_reserve(views, 1.5)
    torch.cuda.synchronize()
    assert_bitexact(views[0][3].cpu().numpy().reshape(v.xres, v.yres, 3), want,
                    "leftover long lists, device counts")
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            R.render_batched(views)
    torch.cuda.synchronize()
    for rep in range(2):
        views[0][3].fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert_bitexact(views[0][3].cpu().numpy().reshape(v.xres, v.yres, 3), want,
                        "leftover long lists, replay %d" % rep)
        assert R.capture_status() == 0
    del g

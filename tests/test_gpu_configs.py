"""GPU parity at the full size of each benchmark workload, through the C-ABI.

  C0  plain: 1M particles (seed 1000), 800^2, additive — bit-exact against the
      oracle and against the genuine host_rendering (oracle/_ref, when built);
  C1  16M particles, 2048^2: 8M gas (ordered, brightness 0.5) + 8M stars
      (additive, hsml 0.05, brightness 2.0) — bit-exact against the oracle in
      the stable (z, index) depth order;
  C2  the 64-view batch (synth.batch_jobs): every batched image equals its
      plain render bit for bit, and one view of each image side (both modes
      among them) equals the oracle;
  C3  the differentiable workload: 2M particles at 512^2 additive and 500k
      ordered, loss sum (pic - target)^2 against the render of the perturbed
      field (synth.perturbed): GPU gradients against the FP64 adjoint and
      chain (orc_adjoint64 + chain64, pinned against FP64 central differences
      on the CPU) on the 200 particles of seed 1303, within 1e-3;
  C4  200M particles at 4096^2, additive over all plus the ordered gas pass:
      two renders are bitwise equal, half the chunk size gives the same image,
      and a 1/100 subsample at 4096^2 is bit-exact against the oracle.
Every size here is the workload's own; each test prints what it runs first.
"""
import math
import sys
import time

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


def gpu_render(R, dev, view, prm):
    import splotch_amd as S
    pic = torch.empty(view.xres * view.yres * 3, dtype=torch.float32, device="cuda")
    R.render(dev, S.camera_setup(view), prm, pic)
    torch.cuda.synchronize()
    out = pic.cpu().numpy().reshape(view.xres, view.yres, 3)
    del pic
    return out


def oracle_two_pass(fld, view, pal, modes, brightness):
    """The oracle's image of a field whose type t is rendered in the passes of
    modes[t] (1 additive, 0 ordered, 2 both): P_ordered + P_additive, the
    ordered pass in stable (z, index) order."""
    p = O.project_colorize(fld, view, [pal] * len(modes), brightness=brightness)
    typ = p.type if fld.type is not None else np.zeros(p.n, np.uint8)
    ordm = np.isin(typ, [t for t, m in enumerate(modes) if m != 1])
    addm = np.isin(typ, [t for t, m in enumerate(modes) if m != 0])
    pic = np.zeros((view.xres, view.yres, 3), np.float32)
    have_ord = bool((p.active.astype(bool) & ordm).any())
    if ordm.any():
        g = p.take(np.arange(p.n))
        g.active = (p.active * ordm).astype(np.uint8)
        g = g.take(O.sort_perm(g, 1, stable=True))
        pic = O.render(g, view.xres, view.yres, False, 0.2, 50)
    if addm.any():
        s = p.take(np.arange(p.n))
        s.active = (p.active * addm).astype(np.uint8)
        a = O.render(s, view.xres, view.yres, True, 0.2, 50)
        pic = (pic + a).astype(np.float32) if have_ord else a
    return pic


def test_c0_plain_1M_800(R):
    import splotch_amd as S
    log("case C0: plummer_field(1M, seed 1000), 800x800, additive")
    fld = synth.plummer_field(1_000_000, 1000)
    view = synth.default_view(800)
    pal = synth.default_palette()
    got = gpu_render(R, S.to_device(fld), view, S.params(a_eq_e=True))
    want, _ = O.host_rendering(fld, view, [pal], a_eq_e=True)
    assert_bitexact(got, want, "C0 vs oracle")
    if O.ref_available():
        gen = O.ref_host_rendering(fld, view, [pal], {"xres": 800, "yres": 800, "fov": 30.0,
                                                      "a_eq_e": True})
        assert_bitexact(got, gen, "C0 vs genuine host_rendering")
        log("C0: bit-exact vs oracle and genuine host_rendering")


def test_c1_gas_stars_16M_2048(R):
    import splotch_amd as S
    log("case C1: gas_and_stars(8M, 8M, 1100, 1101), 2048x2048, gas ordered + stars additive")
    t0 = time.time()
    fld = synth.gas_and_stars(8_000_000, 8_000_000, 1100, 1101)
    view = synth.default_view(2048)
    pal = synth.default_palette()
    prm = S.params(ptypes=2, a_eq_e=[0, 1], brightness=[0.5, 2.0])
    got = gpu_render(R, S.to_device(fld), view, prm)
    log("C1: GPU frame done (%.1f s)" % (time.time() - t0))
    want = oracle_two_pass(fld, view, pal, [0, 1], [0.5, 2.0])
    log("C1: oracle frame done (%.1f s)" % (time.time() - t0))
    assert_bitexact(got, want, "C1 vs oracle (stable depth order)")


def test_c2_batch_64_views(R):
    import splotch_amd as S
    log("case C2: batch_jobs(64): 64 views, 100k-4M particles, 256^2-1024^2, half ordered")
    t0 = time.time()
    jobs = synth.batch_jobs(64)
    views = []
    for bj in jobs:
        views.append((S.to_device(bj.field), S.camera_setup(bj.view),
                      S.params(a_eq_e=bj.additive),
                      torch.empty(bj.view.xres * bj.view.yres * 3, dtype=torch.float32,
                                  device="cuda")))
    log("C2: %d particles generated (%.1f s)" % (sum(b.field.n for b in jobs), time.time() - t0))
    R.render_batched(views)
    torch.cuda.synchronize()
    batched = [v[3].cpu().numpy() for v in views]
    for j, (dev, cam, prm, pic) in enumerate(views):
        R.render(dev, cam, prm, pic)
        torch.cuda.synchronize()
        assert_bitexact(batched[j], pic.cpu().numpy(), "C2 view %d batched vs plain" % j)
    log("C2: 64 batched views equal the plain renders (%.1f s)" % (time.time() - t0))
    # one view of each side vs the oracle, both modes among them
    pal = synth.default_palette()
    chosen = {}
    for j, bj in enumerate(jobs):
        chosen.setdefault(bj.view.xres, j)
    picks = list(chosen.values())
    assert len({jobs[j].additive for j in picks}) == 2, "both modes among the checked views"
    for j in picks:
        bj = jobs[j]
        want, _ = O.host_rendering(bj.field, bj.view, [pal], a_eq_e=bj.additive)
        got = batched[j].reshape(bj.view.xres, bj.view.yres, 3)
        assert_bitexact(got, want, "C2 view %d (%d^2, %s) vs oracle"
                        % (j, bj.view.xres, "additive" if bj.additive else "ordered"))
    log("C2: views %s equal the oracle (%.1f s)" % (picks, time.time() - t0))


def _float_state(p):
    return {"sx": p.x.astype(np.float64), "sy": p.y.astype(np.float64),
            "r": p.r.astype(np.float64), "e0": p.e0.astype(np.float64),
            "e1": p.e1.astype(np.float64), "e2": p.e2.astype(np.float64)}


@pytest.mark.parametrize("name,n,seed,aeq", [("additive_2M", 2_000_000, 1300, True),
                                             ("ordered_500k", 500_000, 1301, False)])
def test_c3_gradients(R, name, n, seed, aeq):
    import splotch_amd as S
    log("case C3 %s: plummer_field(%d, seed %d), 512x512, loss vs perturbed target" %
        (name, n, seed))
    t0 = time.time()
    fld = synth.plummer_field(n, seed)
    tgt = synth.perturbed(fld, 1302)
    view = synth.default_view(512)
    pal = synth.default_palette()
    prm = S.params(a_eq_e=aeq)
    cam = S.camera_setup(view)
    target = torch.empty(512 * 512 * 3, dtype=torch.float32, device="cuda")
    R.render(S.to_device(tgt), cam, prm, target)
    pic = torch.empty_like(target)
    dev = S.to_device(fld)  # the backward reads the forward's particle arrays: keep them
    R.render(dev, cam, prm, pic)
    grad = (2.0 * (pic - target)).contiguous()
    out = R.backward(grad, n, screen=True)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    del dev
    G = grad.cpu().numpy().astype(np.float64).reshape(512, 512, 3)
    log("C3 %s: GPU forward + backward done (%.1f s)" % (name, time.time() - t0))
    p, perm = O.frozen_model(fld, view, [pal], a_eq_e=aeq)
    gro = O.adjoint64(p, _float_state(p), G, 512, 512, a_eq_e=aeq)
    gs = {k: np.zeros(n) for k in gro}
    for k in gro:
        gs[k][perm] = gro[k]
    gw = O.chain64(fld, view, [pal], gs)
    log("C3 %s: FP64 adjoint + chain done (%.1f s)" % (name, time.time() - t0))
    sel = np.random.default_rng(1303).choice(n, size=200, replace=False)
    scr = {k: got["screen"][:, j] for j, k in enumerate(("sx", "sy", "r", "e0", "e1", "e2"))}
    for what, g, ref in (("screen", scr, gs), ("world", got, gw)):
        worst_sel, worst_all = {}, {}
        for k in ref:
            a = np.asarray(g[k], np.float64)
            b = np.asarray(ref[k], np.float64)
            err = np.abs(a - b) / (np.abs(b) + 1e-4 * max(np.abs(b).max(), 1e-300))
            worst_sel[k] = float(err[sel].max())
            worst_all[k] = float(err.max())
        log("C3 %s %s: worst relative error, 200 particles of seed 1303: %s; all %d: %s"
            % (name, what, worst_sel, n, worst_all))
        for k, v in worst_sel.items():
            assert v <= 1e-3, (name, what, k, worst_sel)
        for k, v in worst_all.items():
            assert v <= 1e-3, (name, what, k, worst_all)


@pytest.mark.timeout(600)
def test_c4_large_share_200M_4096(R):
    import splotch_amd as S
    log("case C4: large_share(200M, seed 1400), 4096x4096, additive over all + ordered gas")
    t0 = time.time()
    fld = synth.large_share()
    log("C4: generated %d particles (%.1f s)" % (fld.n, time.time() - t0))
    view = synth.default_view(4096)
    pal = synth.default_palette()
    modes = synth.LARGE_SHARE_MODES
    dev = S.to_device(fld)
    prm = S.params(ptypes=2, a_eq_e=modes, brightness=[0.5, 2.0])
    a = gpu_render(R, dev, view, prm)
    log("C4: frame 1 done (%.1f s)" % (time.time() - t0))
    b = gpu_render(R, dev, view, prm)
    log("C4: frame 2 done (%.1f s)" % (time.time() - t0))
    assert_bitexact(a, b, "C4 repeat")
    _, counts = R.stage_times()
    half = ((1 << 31) - (1 << 20)) // 2
    c = gpu_render(R, dev, view, S.params(ptypes=2, a_eq_e=modes, brightness=[0.5, 2.0],
                                          hip_chunk_entries=half))
    log("C4: half-chunk frame done (%.1f s)" % (time.time() - t0))
    assert_bitexact(c, a, "C4 chunk size halved")
    del dev, a, b, c
    torch.cuda.empty_cache()
    idx = np.sort(np.random.default_rng(7).choice(fld.n, size=fld.n // 100, replace=False))
    sub = synth.Field(fld.x[idx], fld.y[idx], fld.z[idx], fld.hsml[idx], fld.inten[idx],
                      fld.col[idx], fld.type[idx])
    del fld
    got = gpu_render(R, S.to_device(sub), view, prm)
    want = oracle_two_pass(sub, view, pal, modes, [0.5, 2.0])
    log("C4: 1/100 subsample (%d particles) GPU and oracle done (%.1f s)" % (sub.n,
                                                                         time.time() - t0))
    assert_bitexact(got, want, "C4 1/100 subsample vs oracle")

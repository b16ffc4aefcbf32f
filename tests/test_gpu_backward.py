"""GPU parity of the differentiable form (splhip_render_backward), through the C-ABI.

The gradient is that of the smooth model on the forward's frozen footprint
(see include/splotch_hip.h).  Reference: the oracle's FP64 adjoint
(orc_adjoint64, screen space) and FP64 chain (chain64, world space), fed the
same dL/dpic and evaluated at the same point as the GPU — the float32
projected state of the forward render (for a particle of radius ~1 px the
screen gradient is a near-cancelling sum, so moving the evaluation point by a
float32 rounding of its position moves it by ~1e-3 relative); that pair is
itself pinned against FP64 central differences in test_oracle_adjoint.py.  Bar: per parameter, |g_gpu - g_64| <= 1e-3 *
(|g_64| + 1e-4 * max|g_64|).  The autograd test wraps forward and backward in
a torch.autograd.Function (tests only) and checks torch's chain against a
direct call.
"""
import math
import sys

import numpy as np
import pytest

import pyoracle as O
from splotch_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RTOL, FLOOR = 1e-3, 1e-4


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


def _compare(got, ref, what):
    worst = {}
    for k in ref:
        a = np.asarray(got[k], np.float64)
        b = np.asarray(ref[k], np.float64)
        err = np.abs(a - b) / (np.abs(b) + FLOOR * max(np.abs(b).max(), 1e-300))
        worst[k] = float(err.max())
    log(what, "worst relative error:", worst)
    for k, v in worst.items():
        assert v <= RTOL, (what, k, worst)


def _float_state(p):
    """Screen-space state of the float32 forward (render order), in double."""
    return {"sx": p.x.astype(np.float64), "sy": p.y.astype(np.float64),
            "r": p.r.astype(np.float64), "e0": p.e0.astype(np.float64),
            "e1": p.e1.astype(np.float64), "e2": p.e2.astype(np.float64)}


def _oracle_grads(fld, view, pal, G, a_eq_e, sel=None):
    """FP64 screen and world gradients for the particles `sel` (one pass)."""
    sub = fld if sel is None else synth.Field(fld.x[sel], fld.y[sel], fld.z[sel], fld.hsml[sel],
                                             fld.inten[sel], fld.col[sel])
    p, perm = O.frozen_model(sub, view, [pal], a_eq_e=a_eq_e)
    so = _float_state(p)
    gro = O.adjoint64(p, so, G, view.xres, view.yres, a_eq_e=a_eq_e)
    g = {k: np.zeros(sub.n) for k in gro}
    for k in gro:
        g[k][perm] = gro[k]
    return g, O.chain64(sub, view, [pal], g)


def _gpu(R, fld, view, prm, G):
    import splotch_amd as S
    dev = S.to_device(fld)
    cam = S.camera_setup(view)
    pic = torch.empty(view.xres * view.yres * 3, dtype=torch.float32, device="cuda")
    R.render(dev, cam, prm, pic)
    g = torch.from_numpy(G.astype(np.float32).reshape(-1)).cuda()
    out = R.backward(g, fld.n, screen=True)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.parametrize("a_eq_e,n,res,seed", [(True, 60_000, (256, 224), 1300),
                                               (False, 30_000, (192, 192), 1301)])
def test_backward_vs_fp64_adjoint(R, a_eq_e, n, res, seed):
    import splotch_amd as S
    log("case backward", a_eq_e, n, res, seed)
    fld = synth.plummer_field(n, seed, hsml_mu=math.log(0.5))
    view = synth.default_view(*res)
    pal = synth.default_palette()
    G = np.random.default_rng(seed).normal(0.0, 1.0, (res[0], res[1], 3))
    got = _gpu(R, fld, view, S.params(a_eq_e=a_eq_e), G)
    gs, gw = _oracle_grads(fld, view, pal, G, a_eq_e)
    scr = {k: got["screen"][:, j] for j, k in enumerate(("sx", "sy", "r", "e0", "e1", "e2"))}
    _compare(scr, gs, "screen-space a_eq_e=%s" % a_eq_e)
    _compare({k: got[k] for k in gw}, gw, "world-space a_eq_e=%s" % a_eq_e)


@pytest.mark.parametrize("a_eq_e", [False, True])
def test_backward_long_lists(R, a_eq_e):
    """A dense cluster: 8x8 blocks whose lists span many 2048-entry backward
    segments, so the per-segment maps and their join carry the result."""
    import splotch_amd as S
    rng = np.random.default_rng(1350)
    n = 24_000
    xyz = (50.0 + rng.uniform(-5.0, 5.0, (3, n))).astype(np.float32)
    fld = synth.Field(xyz[0], xyz[1], xyz[2], rng.uniform(0.5, 1.5, n).astype(np.float32),
                      rng.lognormal(0.0, 0.5, n).astype(np.float32),
                      rng.uniform(0.0, 1.0, n).astype(np.float32))
    view = synth.View((50.0, 50.0, -150.0), (50.0, 50.0, 50.0), (0.0, 1.0, 0.0), 48, 40, 3.0)
    pal = synth.default_palette()
    G = np.random.default_rng(1351).normal(0.0, 1.0, (48, 40, 3))
    got = _gpu(R, fld, view, S.params(a_eq_e=a_eq_e), G)
    off, idx, _ = R.bins(S.to_device(fld), S.camera_setup(view), S.params(a_eq_e=a_eq_e),
                         ordered=not a_eq_e)
    log("long lists: max tile list", int(np.diff(off).max()))
    gs, gw = _oracle_grads(fld, view, pal, G, a_eq_e)
    scr = {k: got["screen"][:, j] for j, k in enumerate(("sx", "sy", "r", "e0", "e1", "e2"))}
    _compare(scr, gs, "long lists screen a_eq_e=%s" % a_eq_e)
    _compare({k: got[k] for k in gw}, gw, "long lists world a_eq_e=%s" % a_eq_e)


def test_backward_two_types(R):
    """Ordered gas + additive stars in one view: pic = P_ordered + P_additive,
    so each pass receives the same dL/dpic."""
    import splotch_amd as S
    gas = synth.plummer_field(15_000, 1310, ptype=0, hsml_mu=math.log(0.5))
    stars = synth.plummer_field(15_000, 1311, hsml_fixed=0.3, ptype=1)
    fld = synth.concat(gas, stars)
    view = synth.default_view(160, 176)
    pal = synth.default_palette()
    G = np.random.default_rng(1312).normal(0.0, 1.0, (160, 176, 3))
    got = _gpu(R, fld, view, S.params(ptypes=2, a_eq_e=[0, 1]), G)
    ref = {k: np.zeros(fld.n) for k in ("x", "y", "z", "hsml", "inten", "col")}
    for t, aeq in ((0, False), (1, True)):
        sel = np.flatnonzero(fld.type == t)
        _, gw = _oracle_grads(fld, view, pal, G, aeq, sel)
        for k in ref:
            ref[k][sel] = gw[k]
    _compare({k: got[k] for k in ref}, ref, "two types")


def test_backward_autograd_function(R):
    """torch.autograd.Function over the C-ABI forward/backward (tests only)."""
    import splotch_amd as S
    fld = synth.plummer_field(20_000, 1320, hsml_mu=math.log(0.5))
    view = synth.default_view(128, 128)
    cam = S.camera_setup(view)
    prm = S.params(a_eq_e=False)
    keys = ("x", "y", "z", "hsml", "inten", "col")

    class Splat(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, y, z, hsml, inten, col):
            pic = torch.empty(view.xres * view.yres * 3, dtype=torch.float32, device="cuda")
            R.render(dict(x=x, y=y, z=z, hsml=hsml, inten=inten, col=col), cam, prm, pic)
            ctx.n = x.numel()
            return pic

        @staticmethod
        def backward(ctx, gpic):
            g = R.backward(gpic.contiguous(), ctx.n)
            return tuple(g[k] for k in keys)

    dev = S.to_device(fld)
    leaves = [dev[k].clone().requires_grad_(True) for k in keys]
    target = torch.rand(view.xres * view.yres * 3, device="cuda")
    pic = Splat.apply(*leaves)
    loss = ((pic - target) ** 2).sum()
    grads = torch.autograd.grad(loss, leaves)
    # direct call with dL/dpic = 2 (pic - target)
    pic2 = torch.empty_like(pic)
    R.render(dev, cam, prm, pic2)
    direct = R.backward((2.0 * (pic2 - target)).contiguous(), fld.n)
    torch.cuda.synchronize()
    for k, g in zip(keys, grads):
        d = direct[k]
        assert torch.isfinite(g).all()
        err = ((g - d).abs() / (d.abs() + 1e-6 * d.abs().max())).max().item()
        assert err <= 1e-5, (k, err)


def test_backward_vector_colour(R):
    import splotch_amd as S
    fld = synth.plummer_field(20_000, 1330, hsml_mu=math.log(0.5))
    rng = np.random.default_rng(1331)
    fld.c1 = rng.uniform(0, 1, fld.n).astype(np.float32)
    fld.c2 = rng.uniform(0, 1, fld.n).astype(np.float32)
    view = synth.default_view(128, 144)
    pal = synth.default_palette()
    G = rng.normal(0.0, 1.0, (128, 144, 3))
    dev = S.to_device(fld)
    cam = S.camera_setup(view)
    pic = torch.empty(128 * 144 * 3, dtype=torch.float32, device="cuda")
    R.render(dev, cam, S.params(col_vector=[1]), pic)
    out = R.backward(torch.from_numpy(G.astype(np.float32).reshape(-1)).cuda(), fld.n,
                     vector=True)
    torch.cuda.synchronize()
    p, perm = O.frozen_model(fld, view, [pal], col_vector=[1])
    gro = O.adjoint64(p, _float_state(p), G, 128, 144)
    g = {k: np.zeros(fld.n) for k in gro}
    for k in gro:
        g[k][perm] = gro[k]
    gw = O.chain64(fld, view, [pal], g, col_vector=[1])
    _compare({k: out[k].cpu().numpy() for k in gw}, gw, "vector colour")


def test_backward_state_errors(R):
    import splotch_amd as S
    fld = synth.plummer_field(2000, 1340)
    view = synth.default_view(64, 64)
    dev = S.to_device(fld)
    cam = S.camera_setup(view)
    g = torch.zeros(64 * 64 * 3, device="cuda")
    R.bins(dev, cam, S.params(), ordered=False)
    with pytest.raises(S.SplhipError):
        R.backward(g, fld.n)  # bins leave no render to differentiate
    pic = torch.empty(64 * 64 * 3, device="cuda")
    R.render(dev, cam, S.params(hip_region=16), pic)
    with pytest.raises(S.SplhipError):
        R.backward(g, fld.n)  # 16 px regions have no work items
    R.render_batched([(dev, cam, S.params(), pic)])
    with pytest.raises(S.SplhipError):
        R.backward(g, fld.n)  # batches are not differentiated


@pytest.mark.parametrize("a_eq_e", [True, False])
def test_backward_deterministic(R, a_eq_e):
    """The per-particle sums are taken in a fixed order (per-entry partials,
    entries sorted stably by particle, FP64 sums): two backward calls, and
    two forward + backward rounds, give the same gradients bit for bit."""
    import splotch_amd as S
    log("case backward determinism", a_eq_e)
    fld = synth.plummer_field(80_000, 1310, halo_frac=0.9, n_halos=3, hsml_mu=math.log(0.6))
    view = synth.default_view(256, 224)
    G = np.random.default_rng(1311).normal(0.0, 1.0, (256, 224, 3))
    dev = S.to_device(fld)  # the backward reads the forward's particle arrays
    cam = S.camera_setup(view)
    pic = torch.empty(view.xres * view.yres * 3, dtype=torch.float32, device="cuda")
    g = torch.from_numpy(G.astype(np.float32).reshape(-1)).cuda()
    runs = []
    for _ in range(2):
        R.render(dev, cam, S.params(a_eq_e=a_eq_e), pic)
        got = {k: v.cpu().numpy() for k, v in R.backward(g, fld.n, screen=True).items()}
        again = R.backward(g, fld.n, screen=True)
        torch.cuda.synchronize()
        for k, v in again.items():
            assert np.array_equal(v.cpu().numpy().view(np.uint32), got[k].view(np.uint32)), k
        runs.append(got)
    for k in runs[0]:
        assert np.array_equal(runs[0][k].view(np.uint32), runs[1][k].view(np.uint32)), k

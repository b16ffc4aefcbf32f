"""The FP64 adjoint of the oracle (orc_adjoint64 + chain64) against FP64
central differences of the FP64 forward (orc_render64 on the frozen footprint,
screen state from project64), h = 1e-4, on 200 random particles.

Loss L = sum (P64 - target)^2 with a fixed target image.  Bar: per parameter,
|g_adj - g_fd| <= 1e-3 * (|g_fd| + 1e-4 * max|g_fd|) — relative 1e-3 with a
floor of 1e-4 of the largest gradient of that parameter over the sample.
Colour: particles whose colour lies within 2h of a palette node are skipped
(the palette is piecewise linear, so a difference across a node is one-sided).
"""
import math

import numpy as np
import pytest

import pyoracle as O
from splotch_amd import synth

H = 1e-4
PARAMS = ("x", "y", "z", "hsml", "inten", "col")


def _setup(a_eq_e, seed):
    fld = synth.plummer_field(1500, seed, hsml_mu=math.log(1.0), n_halos=4)
    view = synth.default_view(64, 56)
    pal = synth.default_palette()
    p, perm = O.frozen_model(fld, view, [pal], a_eq_e=a_eq_e)
    return fld, view, pal, p, perm


def _loss_and_grad(fld, view, pal, p, perm, a_eq_e, target):
    s = O.project64(fld, view, [pal])
    so = {k: v[perm] for k, v in s.items()}
    P = O.render64(p, so, view.xres, view.yres, a_eq_e=a_eq_e)
    G = 2.0 * (P - target)
    gro = O.adjoint64(p, so, G, view.xres, view.yres, a_eq_e=a_eq_e)
    g = {k: np.zeros(fld.n) for k in gro}
    for k in gro:
        g[k][perm] = gro[k]
    return float(((P - target) ** 2).sum()), O.chain64(fld, view, [pal], g)


def _loss(fld, view, pal, p, perm, a_eq_e, target):
    s = O.project64(fld, view, [pal])
    so = {k: v[perm] for k, v in s.items()}
    P = O.render64(p, so, view.xres, view.yres, a_eq_e=a_eq_e)
    return float(((P - target) ** 2).sum())


def _perturbed(fld, name, i, d):
    arrs = {k: np.array(getattr(fld, k), dtype=np.float64 if k != "type" else None)
            for k in ("x", "y", "z", "hsml", "inten", "col")}
    arrs[name][i] += d
    return synth.Field(arrs["x"], arrs["y"], arrs["z"], arrs["hsml"], arrs["inten"], arrs["col"],
                       fld.type)


@pytest.mark.parametrize("a_eq_e", [True, False])
def test_adjoint64_vs_central_differences(a_eq_e):
    fld, view, pal, p, perm = _setup(a_eq_e, 1303 if a_eq_e else 1304)
    rng = np.random.default_rng(1303)
    s0 = O.project64(fld, view, [pal])
    so = {k: v[perm] for k, v in s0.items()}
    P0 = O.render64(p, so, view.xres, view.yres, a_eq_e=a_eq_e)
    target = 0.5 * P0 + rng.normal(0.0, 0.1 * np.abs(P0).mean(), P0.shape)
    L0, g = _loss_and_grad(fld, view, pal, p, perm, a_eq_e, target)
    # candidates: active particles (in render order p.active) with a footprint
    act = np.zeros(fld.n, bool)
    act[perm] = p.active.astype(bool)
    cand = np.flatnonzero(act & (np.abs(g["x"]) + np.abs(g["inten"]) > 0))
    assert len(cand) >= 200, len(cand)
    sample = rng.choice(cand, 200, replace=False)
    nodes = np.asarray(pal[0], np.float64)
    fd = {k: np.zeros(200) for k in PARAMS}
    ok_col = np.ones(200, bool)
    for j, i in enumerate(sample):
        for name in PARAMS:
            if name == "col" and np.min(np.abs(nodes - float(fld.col[i]))) < 2 * H:
                ok_col[j] = False
                continue
            lp = _loss(_perturbed(fld, name, i, +H), view, pal, p, perm, a_eq_e, target)
            lm = _loss(_perturbed(fld, name, i, -H), view, pal, p, perm, a_eq_e, target)
            fd[name][j] = (lp - lm) / (2 * H)
    worst = {}
    for name in PARAMS:
        a, b = g[name][sample], fd[name]
        if name == "col":
            a, b = a[ok_col], b[ok_col]
        err = np.abs(a - b) / (np.abs(b) + 1e-4 * np.abs(b).max())
        worst[name] = float(err.max())
        assert err.max() <= 1e-3, (name, worst)
    print("FP64 adjoint vs FD, a_eq_e=%s, worst relative error per parameter: %s" % (a_eq_e, worst))

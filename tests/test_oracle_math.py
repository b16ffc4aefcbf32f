"""Closed forms, invariants and brute force that pin the oracle independently of
the fixtures (each chosen so a dropped term, wrong sign/index or transposed
operand fails)."""
import math

import numpy as np
import pytest

import pyoracle as O
from splotch_amd import synth


def test_exptable_nodes_and_bounds():
    t1, t2, expfac, tl = O.exptable()
    # tables are expf at the node points (splotchutils.h:193-197)
    m = np.arange(1024)
    a1 = (m * 1024 / np.float32(expfac)).astype(np.float32)
    a2 = (m / np.float32(expfac)).astype(np.float32)
    for t, a in ((t1, a1), (t2, a2)):
        ref = np.exp(a.astype(np.float64)).astype(np.float32)   # correctly rounded expf
        ulp = np.abs(t.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
        assert ulp.max() <= 1
    assert np.float32(expfac) == np.float32(1048576 / np.float32(-20.0))
    assert abs(tl - math.sqrt(2 * 20 / 1048576)) < 1e-9
    x = -np.linspace(0, 19.99, 200001).astype(np.float32)
    op = O.exp_op(x)
    rel = np.abs(op / np.exp(x.astype(np.float64)) - 1.0)
    assert rel.max() <= 20.0 / 2 ** 20 * 1.0001 + 1e-6   # one truncation step of 1.907e-5
    assert O.exp_op(np.array([0.0], np.float32))[0] == 1.0
    assert O.exp_op(np.array([0.5, 3.0], np.float32)).tolist() == [1.0, 1.0]   # negative index
    assert O.exp_op(np.array([-20.0, -25.0], np.float32)).tolist() == [0.0, 0.0]
    small = np.array([-1e-3, 2e-3, -6e-3], np.float32)
    assert np.array_equal(O.exp_op(small, True), small)   # Taylor branch returns arg


def test_camera_closed_form():
    v = synth.View((10.0, -40.0, 70.0), (55.0, 45.0, 40.0), (0.2, 0.1, 1.0), 300, 200, 45.0)
    cam = O.camera(v)
    M = np.array(cam.mat[:], np.float64).reshape(3, 4)
    R = M[:, :3]
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)            # orthonormal
    look = M @ np.array([55.0, 45.0, 40.0, 1.0])
    assert look[0] == pytest.approx(0, abs=1e-4) and look[1] == pytest.approx(0, abs=1e-4)
    assert -look[2] == pytest.approx(np.linalg.norm(np.subtract(v.lookat, v.campos)), rel=1e-6)
    cam0 = M @ np.array(list(v.campos) + [1.0])
    assert np.allclose(cam0, 0, atol=1e-4)
    assert cam.xfac == np.float32(np.float32(150) / (np.float32(math.tan(22.5 * math.pi / 180))
                                                     * np.float32(cam.dist)))
    with pytest.raises(ValueError):   # sky parallel to the view axis
        O.camera(synth.View((50, 50, -150), (50, 50, 50), (0, 0, 1), 64, 64))


def _one(x, y, r, e, n_extra=0):
    p = O.Projected(1 + n_extra)
    p.x[0], p.y[0], p.r[0] = x, y, r
    p.e0[0], p.e1[0], p.e2[0] = e
    p.active[0] = 1
    return p


def test_single_particle_known_answer():
    # pixel value = -e * tab(stp*dy*dy) * tab(stp*dx^2) inside the footprint
    xres, yres = 40, 30
    p = _one(17.3, 11.6, 6.0, (1.0, 0.5, 0.25))
    pic = O.render(p, xres, yres, True, 0.2, 50)
    h2s = np.float32(0.5 * math.pi ** (-1 / 6))
    sigma = np.float32(h2s * np.float32(6.0))
    stp = np.float32(np.float32(-1.0) / (sigma * sigma))
    rf = np.float32(6.0) * np.float32(0.75)
    X, Y = 17, 12
    dx = np.float32(X) - np.float32(17.3)
    dy = np.float32(Y) - np.float32(11.6)
    att = O.exp_op(np.array([stp * dy * dy], np.float32))[0] * \
        O.exp_op(np.array([stp * (dx * dx)], np.float32))[0]
    assert pic[X, Y, 0] == np.float32(-1.0) * np.float32(att)
    assert pic[X, Y, 1] == np.float32(-0.5) * np.float32(att)
    # outside the circle of radius 0.75 r nothing is written
    assert pic[17 + 6, 12, 0] == 0 and pic[17, 12 + 6, 0] == 0
    assert (pic[:, :, 0] <= 0).all()
    # footprint bound: |X - x| <= rfacr
    xs, ys = np.nonzero(pic[:, :, 0])
    assert (np.abs(xs - 17.3) <= rf + 1e-5).all() and (np.abs(ys - 11.6) <= rf + 1e-5).all()


def test_two_particle_ordered_closed_form():
    # P1 = q1 + (0 - q1)*T1, P2 = q2 + (P1 - q2)*T2 with T = 1 + expm1(-e*att)
    p = O.Projected(2)
    p.x[:] = [10.2, 10.4]; p.y[:] = [10.1, 9.8]; p.r[:] = [5.0, 4.0]
    p.e0[:] = [0.8, 0.3]; p.e1[:] = [0.4, 0.9]; p.e2[:] = [0.1, 0.2]; p.active[:] = 1
    pic = O.render(p, 24, 24, False, 0.2, 50)
    X, Y = 10, 10
    vals = []
    for i in range(2):
        h2s = np.float32(0.5 * math.pi ** (-1 / 6))
        sig = np.float32(h2s * p.r[i])
        stp = np.float32(np.float32(-1) / (sig * sig))
        dx = np.float32(X) - p.x[i]
        dy = np.float32(Y) - p.y[i]
        att = np.float32(O.exp_op(np.array([stp * dy * dy], np.float32))[0] *
                         O.exp_op(np.array([stp * (dx * dx)], np.float32))[0])
        vals.append(att)
    P = np.zeros(3, np.float32)
    for i in range(2):
        e = np.array([p.e0[i], p.e1[i], p.e2[i]], np.float32)
        q = (e / (e + np.float32(0.2))).astype(np.float32)
        a = -e
        att = vals[i]
        if np.float32(np.abs(a).max() * att) < np.float32(O.exptable()[3]):
            P = (P + ((P - q) * a) * att).astype(np.float32)
        else:
            P = (P + O.exp_op((att * a).astype(np.float32), True) * (P - q)).astype(np.float32)
    assert np.array_equal(pic[X, Y], P)


def test_bruteforce_equals_binned():
    """All-particles-per-pixel loop == binned render, bit for bit (pins that the
    tile lists are conservative and in ascending order)."""
    f = synth.plummer_field(3000, 21)
    v = synth.default_view(90, 70)
    p = O.project_colorize(f, v, [synth.default_palette()])
    for T in (50, 16, 7):
        a = O.render(p, v.xres, v.yres, True, 0.2, T)
        b = O.render_bruteforce(p, v.xres, v.yres, T)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), T


def test_bins_bruteforce():
    f = synth.plummer_field(2000, 22)
    v = synth.default_view(100, 80)
    p = O.project_colorize(f, v, [synth.default_palette()])
    T = 16
    off, idx = O.bins(p, v.xres, v.yres, T)
    ncx, ncy = (100 + T - 1) // T, (80 + T - 1) // T
    rcell = np.float32(math.sqrt(2)) * np.float32(T * 0.5 - 0.5)
    # brute force: every tile x every particle with the circle rule
    for t in range(ncx * ncy):
        ix, iy = t % ncx, t // ncx
        cx = np.float32(0.5 * (T - 1)) + np.float32(ix * T)
        cy = np.float32(0.5 * (T - 1)) + np.float32(iy * T)
        rf = (np.float32(0.75) * p.r).astype(np.float32)
        lo_x = np.maximum(0, np.trunc(p.x - rf + np.float32(1)).astype(np.int64) // T)
        hi_x = np.minimum(ncx - 1, np.trunc(p.x + rf).astype(np.int64) // T)
        lo_y = np.maximum(0, np.trunc(p.y - rf + np.float32(1)).astype(np.int64) // T)
        hi_y = np.minimum(ncy - 1, np.trunc(p.y + rf).astype(np.int64) // T)
        d2 = ((p.x - cx) * (p.x - cx) + (p.y - cy) * (p.y - cy)).astype(np.float32)
        ok = (p.active == 1) & (lo_x <= ix) & (ix <= hi_x) & (lo_y <= iy) & (iy <= hi_y) & \
            (d2 < (rcell + rf) * (rcell + rf))
        assert np.array_equal(np.nonzero(ok)[0].astype(np.uint32), idx[off[t]:off[t + 1]]), t


def test_additivity_and_fp64_twin():
    f = synth.plummer_field(4000, 23)
    v = synth.default_view(80, 80)
    p = O.project_colorize(f, v, [synth.default_palette()])
    half = np.arange(p.n) < p.n // 2
    pa = p.take(np.arange(p.n)); pa.active = (p.active * half).astype(np.uint8)
    pb = p.take(np.arange(p.n)); pb.active = (p.active * ~half).astype(np.uint8)
    full = O.render(p, 80, 80, True, 0.2, 50).astype(np.float64)
    parts = O.render(pa, 80, 80, True, 0.2, 50).astype(np.float64) + \
        O.render(pb, 80, 80, True, 0.2, 50).astype(np.float64)
    assert np.allclose(full, parts, rtol=1e-5, atol=1e-6 * np.abs(full).max())
    s = dict(sx=p.x, sy=p.y, r=p.r, e0=p.e0, e1=p.e1, e2=p.e2)
    p64 = O.render64(p, s, 80, 80, True, 0.2, 50)
    rel = np.abs(full - p64) / np.maximum(np.abs(p64), 1e-6 * np.abs(p64).max())
    # exptable: <= 2 truncation steps per contribution, plus float accumulation
    assert rel.max() < 5e-5


def test_fragments_count():
    """orc_fragments (the count of render_new's per-pixel updates, for the
    fragment rate) equals a numpy count of the same footprint rule over the
    tile lists."""
    f = synth.plummer_field(1500, 24, hsml_mu=math.log(0.6))
    v = synth.default_view(96, 80)
    p = O.project_colorize(f, v, [synth.default_palette()])
    for T in (50, 16):
        off, idx = O.bins(p, v.xres, v.yres, T)
        ncx = (v.xres + T - 1) // T
        want = 0
        for t in range(len(off) - 1):
            x0s, y0s = (t % ncx) * T, (t // ncx) * T
            x1 = min(x0s + T, v.xres) - x0s
            y1 = min(y0s + T, v.yres) - y0s
            for i in idx[off[t]:off[t + 1]]:
                rf = np.float32(p.r[i]) * np.float32(0.75)
                px = np.float32(p.x[i]) - np.float32(x0s)
                py = np.float32(p.y[i]) - np.float32(y0s)
                minx = max(int(np.float32(px - rf + np.float32(1))), 0)
                maxx = min(int(np.float32(px + rf + np.float32(1))), x1)
                miny = max(int(np.float32(py - rf + np.float32(1))), 0)
                maxy = min(int(np.float32(py + rf + np.float32(1))), y1)
                for xx in range(minx, maxx):
                    dxsq = np.float32(np.float32(xx) - px) * np.float32(np.float32(xx) - px)
                    dy = np.sqrt(np.float32(rf * rf - dxsq), dtype=np.float32)
                    lo = max(miny, int(np.float32(py - dy + np.float32(1))))
                    hi = min(maxy, int(np.float32(py + dy + np.float32(1))))
                    want += max(hi - lo, 0)
        assert O.fragments(p, v.xres, v.yres, T) == want, T

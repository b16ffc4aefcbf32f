"""The block cull of region binning never drops a block that render_new
updates.

Binning through regions keeps, of a region's 8x8 blocks, only those whose
pixel square [8 b, 8 b + 7] per axis lies within rfacr + 2 px (and a 2^-16
relative margin) of the particle's tile-frame centre (`region_reach` in
splotch_amd/csrc/kernels.hip).  This checks that criterion, in float32 as
the kernel evaluates it, against the footprint that render_new actually
walks (splotch_host.cc:591-607: columns [int(posx - rfacr + 1),
int(posx + rfacr + 1)) clipped to the tile, and per column the rows
[int(posy - dy + 1), int(posy + dy + 1)) with dy = sqrtf(radsq - dxsq)),
over radii from a tenth of a pixel to 30000 px, where the cancellation in
radsq - dxsq is largest, and centres inside, at the edges of and far outside
the tile.  A criterion half a pixel tighter than the footprint fails both
tests; on these cases even one with no margin passes (the float rounding of
the footprint bounds stays well below a pixel), so the 2 px margin is slack,
not a tuned bound."""
import numpy as np

T = 50          # tile size (default tile_size_cpu)
NB = (T + 7) // 8


def _trunc_i32(v):
    """x86 cvttss2si: truncation toward zero, INT_MIN for NaN / out of range."""
    v = np.asarray(v, np.float32)
    out = np.full(v.shape, -2 ** 31, np.int64)
    ok = np.isfinite(v) & (np.abs(v) < 2.0 ** 31)
    out[ok] = np.trunc(v[ok]).astype(np.int64)
    return out


def _reached_blocks(px, py, rf):
    """(n, NB, NB) bool: block (bx, by) has a pixel that render_new updates."""
    f = np.float32
    n = px.shape[0]
    w = h = T
    radsq = (rf * rf).astype(f)
    minx = np.maximum(_trunc_i32(px - rf + f(1)), 0)
    maxx = np.minimum(_trunc_i32(px + rf + f(1)), w)
    miny = np.maximum(_trunc_i32(py - rf + f(1)), 0)
    maxy = np.minimum(_trunc_i32(py + rf + f(1)), h)
    out = np.zeros((n, NB, NB), bool)
    ys = np.arange(h)
    for x in range(w):
        dx = (f(x) - px).astype(f)
        dxsq = (dx * dx).astype(f)
        with np.errstate(invalid="ignore"):
            dy = np.sqrt((radsq - dxsq).astype(f)).astype(f)
        lo = np.maximum(miny, _trunc_i32(py - dy + f(1)))
        hi = np.minimum(maxy, _trunc_i32(py + dy + f(1)))
        col = (x >= minx) & (x < maxx)
        rows = col[:, None] & (ys[None, :] >= lo[:, None]) & (ys[None, :] < hi[:, None])
        for by in range(NB):
            out[:, x // 8, by] |= rows[:, 8 * by:8 * by + 8].any(axis=1)
    return out


def _kept_blocks(px, py, rf):
    """(n, NB, NB) bool: the cull's criterion, float32 as in the kernel."""
    f = np.float32
    lim = (rf + f(2)).astype(f)
    thr = ((lim * lim).astype(f) * f(1.0000153)).astype(f)
    b0 = (8 * np.arange(NB)).astype(f)
    dxx = np.maximum(np.maximum(b0[None, :] - px[:, None], px[:, None] - (b0[None, :] + f(7))),
                     f(0)).astype(f)
    dyy = np.maximum(np.maximum(b0[None, :] - py[:, None], py[:, None] - (b0[None, :] + f(7))),
                     f(0)).astype(f)
    d2 = ((dxx * dxx).astype(f)[:, :, None] + (dyy * dyy).astype(f)[:, None, :]).astype(f)
    return ~(d2 > thr[:, None, None])


def test_cull_keeps_every_reached_block():
    rng = np.random.default_rng(2024)
    n = 6000
    rf = (10.0 ** rng.uniform(-1, np.log10(30000.0), n)).astype(np.float32)
    # centres: inside the tile, near its edges, and out to rfacr + 10 px away
    span = rng.uniform(-1.0, 1.0, (n, 2)) * (rf[:, None] + 10.0)
    pick = rng.integers(0, 3, n)
    centre = np.where(pick[:, None] == 0, rng.uniform(0, T, (n, 2)),
                      np.where(pick[:, None] == 1, T * rng.integers(0, 2, (n, 2)) + span * 0.05,
                               T / 2 + span))
    px = centre[:, 0].astype(np.float32)
    py = centre[:, 1].astype(np.float32)
    reached = _reached_blocks(px, py, rf)
    kept = _kept_blocks(px, py, rf)
    dropped = reached & ~kept
    assert not dropped.any(), "a reached block is culled: particle %d" % np.argwhere(dropped)[0][0]
    # the cull does drop blocks (it is not vacuous) and keeps few unreached ones
    assert (kept & ~reached).sum() < 0.2 * kept.sum()
    assert (~kept).sum() > 0.05 * kept.size


def test_cull_boundary_cases():
    """Centres placed so the footprint edge lies just inside a block's nearest
    pixel: that block is reached and kept."""
    f = np.float32
    rf = np.array([0.5, 3.0, 17.25, 400.0, 9000.0, 25000.0], f)
    # block (2, 3) covers columns 16..23, rows 24..31: the centre sits on row
    # 27, rf - eps left of column 16
    eps = np.maximum(rf * f(1e-6), f(1e-3)).astype(f)
    px = (f(16) - rf + eps).astype(f)
    py = np.full_like(px, f(27))
    reached = _reached_blocks(px, py, rf)
    kept = _kept_blocks(px, py, rf)
    assert reached[:, 2, 3].all()
    assert not (reached & ~kept).any()

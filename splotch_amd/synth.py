"""Seeded synthetic particle fields and palettes for the render workloads.

The fields stand in for post-``particle_normalize`` simulation data at the
render boundary (splotch/scenemaker.cc:1097 hands the renderer positions,
smoothing lengths, intensities and a colour scalar in [0,1]).  Layout is SoA
float32, the layout the HIP kernels stream.

Generator (all draws from ``numpy.random.default_rng(seed)``, so a seed gives
the same field on any host with this numpy):
  * ``halo_frac`` of the particles in ``n_halos`` Plummer spheres: centres
    U[0,100]^3, scale radii a = exp(U[ln .2, ln 2]), halo masses from
    dN/dM ~ M^-1.9 on [1, 1e3] (inverse CDF), particles per halo by the
    largest-remainder rule on M; radius r = a / sqrt(u^(-2/3) - 1) capped at
    20a, isotropic direction;
  * the rest uniform in [0,100]^3;
  * hsml ~ LogNormal(hsml_mu, hsml_sigma) (or a fixed value), I ~ LogNormal(0,1),
    colour ~ U[0,1].
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Field:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    hsml: np.ndarray
    inten: np.ndarray
    col: np.ndarray
    type: np.ndarray | None = None
    c1: np.ndarray | None = None
    c2: np.ndarray | None = None

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def arrays(self):
        return dict(x=self.x, y=self.y, z=self.z, hsml=self.hsml, inten=self.inten,
                    col=self.col, type=self.type, c1=self.c1, c2=self.c2)


def _largest_remainder(weights: np.ndarray, total: int) -> np.ndarray:
    w = weights / weights.sum()
    raw = w * total
    base = np.floor(raw).astype(np.int64)
    rem = total - int(base.sum())
    if rem > 0:
        order = np.argsort(-(raw - base), kind="stable")
        base[order[:rem]] += 1
    return base


def plummer_field(n: int, seed: int, halo_frac: float = 0.7, n_halos: int = 50,
                  hsml_mu: float = math.log(0.3), hsml_sigma: float = 0.6,
                  hsml_fixed: float | None = None, box: float = 100.0,
                  ptype: int | None = None) -> Field:
    """Particle field of the plain 1M / 800^2 workload family (see module doc)."""
    rng = np.random.default_rng(seed)
    n_halo = int(round(n * halo_frac))
    n_bg = n - n_halo
    centres = rng.uniform(0.0, box, size=(n_halos, 3))
    scale = np.exp(rng.uniform(math.log(0.2), math.log(2.0), size=n_halos))
    # dN/dM ~ M^-1.9 on [1, 1e3]: CDF ~ 1 - M^-0.9 (normalised); inverse CDF
    u = rng.uniform(0.0, 1.0, size=n_halos)
    lo, hi, k = 1.0, 1e3, -0.9
    mass = (lo ** k + u * (hi ** k - lo ** k)) ** (1.0 / k)
    counts = _largest_remainder(mass, n_halo)
    hid = np.repeat(np.arange(n_halos), counts)
    uu = rng.uniform(1e-12, 1.0, size=n_halo)
    rad = scale[hid] / np.sqrt(uu ** (-2.0 / 3.0) - 1.0)
    rad = np.minimum(rad, 20.0 * scale[hid])
    cth = rng.uniform(-1.0, 1.0, size=n_halo)
    phi = rng.uniform(0.0, 2.0 * math.pi, size=n_halo)
    sth = np.sqrt(1.0 - cth * cth)
    pos_h = centres[hid] + rad[:, None] * np.stack([sth * np.cos(phi), sth * np.sin(phi), cth], 1)
    pos_b = rng.uniform(0.0, box, size=(n_bg, 3))
    pos = np.concatenate([pos_h, pos_b], 0)
    if hsml_fixed is not None:
        hsml = np.full(n, hsml_fixed)
    else:
        hsml = rng.lognormal(hsml_mu, hsml_sigma, size=n)
    inten = rng.lognormal(0.0, 1.0, size=n)
    col = rng.uniform(0.0, 1.0, size=n)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    typ = None if ptype is None else np.full(n, ptype, dtype=np.uint8)
    return Field(f32(pos[:, 0]), f32(pos[:, 1]), f32(pos[:, 2]), f32(hsml), f32(inten),
                 f32(col), typ)


def concat(a: Field, b: Field) -> Field:
    def cat(u, v):
        if u is None and v is None:
            return None
        return np.concatenate([u, v])
    ta = a.type if a.type is not None else np.zeros(a.n, np.uint8)
    tb = b.type if b.type is not None else np.zeros(b.n, np.uint8)
    return Field(cat(a.x, b.x), cat(a.y, b.y), cat(a.z, b.z), cat(a.hsml, b.hsml),
                 cat(a.inten, b.inten), cat(a.col, b.col), cat(ta, tb))


def gas_and_stars(n_gas: int = 8_000_000, n_stars: int = 8_000_000, seed_gas: int = 1100,
                  seed_stars: int = 1101) -> Field:
    """Two-type field: gas (type 0, hsml LogN(ln .5, .6)) + stars (type 1, hsml .05)."""
    gas = plummer_field(n_gas, seed_gas, hsml_mu=math.log(0.5), ptype=0)
    stars = plummer_field(n_stars, seed_stars, hsml_fixed=0.05, ptype=1)
    return concat(gas, stars)


# ---------------------------------------------------------------- palettes

def hot_palette(n: int = 256) -> np.ndarray:
    """A 'hot' black-red-yellow-white ramp with integer 0-255 entries (n x 3)."""
    t = np.arange(n) / (n - 1)
    r = np.clip(3.0 * t, 0, 1)
    g = np.clip(3.0 * t - 1.0, 0, 1)
    b = np.clip(3.0 * t - 2.0, 0, 1)
    return np.rint(np.stack([r, g, b], 1) * 255).astype(np.int32)


def palette_text(rgb255: np.ndarray, name: str = "hot") -> str:
    """The .pal text format read by splotch/splotchutils.cc:105-123."""
    lines = [name, "0100", str(len(rgb255))]
    lines += ["%03d %03d %03d" % tuple(int(v) for v in c) for c in rgb255]
    return "\n".join(lines) + "\n"


def parse_palette(text: str):
    """Parse .pal text like splotch/splotchutils.cc:105-123.

    Returns (x, rgb): x[i] = i * (1/(n-1)) as double, rgb[i] = byte/255 as float32.
    """
    tok = text.split()
    n = int(tok[2])
    vals = np.array([float(v) for v in tok[3:3 + 3 * n]], dtype=np.float32).reshape(n, 3)
    step = 1.0 / (n - 1)
    x = np.array([i * step for i in range(n)], dtype=np.float64)
    rgb = (vals / np.float32(255)).astype(np.float32)
    return x, rgb


def default_palette():
    return parse_palette(palette_text(hot_palette(256)))


# ---------------------------------------------------------------- cameras

@dataclass
class View:
    campos: tuple
    lookat: tuple
    sky: tuple
    xres: int
    yres: int
    fov: float = 30.0
    centerpos: tuple | None = None

    @property
    def center(self):
        return self.centerpos if self.centerpos is not None else self.campos


def default_view(xres: int = 800, yres: int | None = None) -> View:
    """Camera (50,50,-150) looking at the box centre, sky (0,1,0), fov 30."""
    return View((50.0, 50.0, -150.0), (50.0, 50.0, 50.0), (0.0, 1.0, 0.0), xres,
                yres if yres is not None else xres, 30.0)


def orbit_view(az: float, el: float, radius: float = 150.0, xres: int = 512,
               yres: int | None = None) -> View:
    c = (50.0 + radius * math.cos(el) * math.cos(az),
         50.0 + radius * math.cos(el) * math.sin(az),
         50.0 + radius * math.sin(el))
    return View(c, (50.0, 50.0, 50.0), (0.0, 0.0, 1.0), xres,
                yres if yres is not None else xres, 30.0)


@dataclass
class BatchJob:
    """One view of the batched workload: its particles, camera and mode."""
    field: Field
    view: View
    additive: bool


def batch_jobs(n_jobs: int = 64, seed: int = 1200, n_scale: float = 1.0,
               n_min: float = 1e5, n_max: float = 4e6) -> list:
    """Batched workload: job j has seed+j particles N_j = round(exp(U[ln n_min,
    ln n_max]) * n_scale) from the plain workload's generator, a square image of
    side 256/512/768/1024 and a camera orbiting the box at radius 150 (azimuth
    U[0, 2pi), elevation asin(U[-0.98, 0.98]), sky +z); even jobs are
    additive, odd jobs ordered (depth-sorted)."""
    out = []
    for j in range(n_jobs):
        rng = np.random.default_rng(seed + j)
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max))) * n_scale))
        side = int(rng.choice([256, 512, 768, 1024]))
        az = rng.uniform(0.0, 2.0 * math.pi)
        el = math.asin(rng.uniform(-0.98, 0.98))
        f = plummer_field(max(n, 1), seed=seed + j)
        out.append(BatchJob(f, orbit_view(az, el, 150.0, side), j % 2 == 0))
    return out


def large_share(n: int = 200_000_000, seed: int = 1400) -> Field:
    """One GPU's share of a large simulation: n particles, half in 5000 Plummer
    halos and half uniform background, hsml log-normal (mu = ln 0.2, sigma =
    1.2).  The halo half is the gas (type 0) and the background half type 1.
    The frame (LARGE_SHARE_MODES) is an additive pass over all particles plus
    a depth-sorted emission-absorption pass over the gas: type 0 takes part in
    both passes (a_eq_e 2), type 1 in the additive one."""
    f = plummer_field(n, seed, halo_frac=0.5, n_halos=5000, hsml_mu=math.log(0.2),
                      hsml_sigma=1.2)
    n_halo = int(round(n * 0.5))
    typ = np.ones(n, dtype=np.uint8)
    typ[:n_halo] = 0
    f.type = typ
    return f


# a_eq_e per type of the large-share frame: gas in both passes, the rest additive
LARGE_SHARE_MODES = [2, 1]


def perturbed(field: Field, seed: int = 1302, sigma: float = 0.05) -> Field:
    """The differentiable workload's target field: positions + N(0, sigma) per
    axis, intensities * U(0.8, 1.2); hsml and colour unchanged."""
    rng = np.random.default_rng(seed)
    n = field.n
    x, y, z = (getattr(field, k) + rng.normal(0, sigma, n).astype(np.float32)
               for k in ("x", "y", "z"))
    inten = (field.inten * rng.uniform(0.8, 1.2, n)).astype(np.float32)
    return Field(np.ascontiguousarray(x, np.float32), np.ascontiguousarray(y, np.float32),
                 np.ascontiguousarray(z, np.float32), field.hsml, inten, field.col, field.type)

"""Synthetic particle and halo sets of the benchmark configurations.

uniform_points: FP32 coordinates uniform in [0, 1000)^3.

clustered: particles drawn from isotropic Gaussian clumps plus a uniform
background, and halos placed at clump centres plus a uniform share -- the
uneven-work configuration (200M particles, 70% in 200k clumps with sigma
uniform in [0.3, 2.0] and power-law sizes of exponent -2 truncated to
[10, 1e5], rescaled so the clump members total 70% of the particles; 5M
halos, 80% at clump centres).  Coordinates are wrapped into the periodic box.

gadget_snapshot: a compressed (Gadget) snapshot of a side^3 grid whose
particles carry uniform uint16 displacements over [-disp, disp].
"""
from __future__ import annotations

import numpy as np

BOX = np.float32(1000.0)
_TOP = np.nextafter(np.float32(1000.0), np.float32(0))


def uniform_points(rng: np.random.Generator, n: int) -> np.ndarray:
    pts = rng.random((n, 3), dtype=np.float32) * BOX
    # float32 rounding of u*1000 can give exactly 1000.0: keep [0, 1000)
    np.minimum(pts, _TOP, out=pts)
    return pts


def _wrap(pts: np.ndarray) -> None:
    np.mod(pts, BOX, out=pts)
    np.minimum(pts, _TOP, out=pts)  # a tiny negative value wraps to 1000.0


def clump_sizes(rng: np.random.Generator, nclumps: int, members: int,
                lo: float = 10.0, hi: float = 1e5) -> np.ndarray:
    """Power-law sizes p(s) ~ s^-2 on [lo, hi] (inverse CDF), rescaled to sum
    to exactly `members`."""
    u = rng.random(nclumps)
    s = 1.0 / (1.0 / lo - u * (1.0 / lo - 1.0 / hi))
    s *= members / s.sum()
    sizes = np.floor(s).astype(np.int64)
    short = members - int(sizes.sum())
    if short > 0:
        sizes[np.argsort(s - sizes)[::-1][:short]] += 1
    return sizes


def clustered(rng: np.random.Generator, n_particles: int, n_halos: int,
              nclumps: int = 200_000, clump_frac: float = 0.7,
              halo_centre_frac: float = 0.8, sigma=(0.3, 2.0)):
    """Returns (particles (n, 3) float32, halos (m, 3) float32)."""
    members = int(round(clump_frac * n_particles))
    centres = uniform_points(rng, nclumps)
    sig = rng.uniform(sigma[0], sigma[1], nclumps).astype(np.float32)
    sizes = clump_sizes(rng, nclumps, members)
    parts = np.empty((n_particles, 3), np.float32)
    which = np.repeat(np.arange(nclumps, dtype=np.int64), sizes)
    step = 1 << 24
    for a in range(0, members, step):
        b = min(members, a + step)
        w = which[a:b]
        parts[a:b] = centres[w] + rng.standard_normal((b - a, 3), dtype=np.float32) * sig[w, None]
    del which
    parts[members:] = uniform_points(rng, n_particles - members)
    _wrap(parts)
    nh_c = int(round(halo_centre_frac * n_halos))
    halos = np.empty((n_halos, 3), np.float32)
    halos[:nh_c] = centres[rng.integers(0, nclumps, nh_c)]
    halos[nh_c:] = uniform_points(rng, n_halos - nh_c)
    return parts, halos


def gadget_snapshot(rng: np.random.Generator, side: int, disp: float = 20.0, chunk: int = 1 << 24):
    """(header bytes, uint16 triplets [side^3][3]) of one file holding the
    whole grid, ids 1..side^3."""
    from . import gadget_header
    n = side ** 3
    uv = np.empty((n, 3), np.uint16)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        uv[a:b] = rng.integers(0, 65536, (b - a, 3), dtype=np.uint16)
    return gadget_header(n, n, 1000.0, -disp, disp, 16, 1), uv

"""Seeded random workloads and engine settings against the oracle.

Each case draws particles (uniform, with exact 0 and 1000 coordinates, or
in Gaussian clumps), halos (uniform with a share on the faces, some
repeated to form tiles, some placed on particles), 1 to 96 shell thresholds
(ascending, or shuffled: first match as the reference's scan,
src/objects/ArraySTR/MultiRangeCount.c:107-113) with r_max up to 600 (a
query box wider than half the box: all eight periodic images,
MultiRangeCount.c:31-71), the particles in one to three files, and a random
combination of the engine's options (lanes per halo, segment, kernel, cell
edge, x split, sort kind, tiles, tile level, halos per tile, LDS staging,
graph replay, overlap, pieces, fold limit) -- then checks every count
against the brute-force oracle.  Work per case is bounded (halos x
particles <= 1.2e8) so the oracle takes at most a few seconds."""
import numpy as np
import pytest

import sjcs_testlib as t

pytestmark = pytest.mark.gpu


def draw_case(seed):
    rng = np.random.default_rng(70_000 + seed)
    npart = int(10 ** rng.uniform(3.0, 5.3))
    rmax = float(10 ** rng.uniform(-0.3, np.log10(600.0)))
    cap = 1.2e8 if rmax < 100 else 4e6
    nh = int(min(10 ** rng.uniform(1.5, 3.6), cap / npart))
    nh = max(nh, 3)
    if rng.random() < 0.3:  # clumped particles
        nc = int(rng.integers(3, 40))
        centres = rng.random((nc, 3)) * 1000.0
        which = rng.integers(0, nc, npart)
        p = centres[which] + rng.standard_normal((npart, 3)) * rng.uniform(0.3, 3.0)
        p = np.mod(p, 1000.0).astype(np.float32)
        p = np.minimum(p, np.nextafter(np.float32(1000.0), np.float32(0)))
    else:
        p = (rng.random((npart, 3), dtype=np.float32) * np.float32(1000.0))
    p[:8] = 1000.0
    p[8:16] = 0.0
    h = rng.random((nh, 3), dtype=np.float32) * np.float32(1000.0)
    m = np.nonzero(rng.random(nh) < 0.15)[0]
    ax = rng.integers(0, 3, m.size)
    lo = rng.random(m.size) < 0.5
    h[m, ax] = np.where(lo, rng.uniform(0, 5, m.size), rng.uniform(995, 1000, m.size))
    if rng.random() < 0.4:  # repeated centres: tiles of identical halos
        k = max(1, nh // 5)
        h[rng.integers(0, nh, k)] = h[rng.integers(0, max(1, nh // 20), k)]
    if rng.random() < 0.3:  # centres on particles
        k = max(1, nh // 10)
        h[rng.integers(0, nh, k)] = p[rng.integers(0, npart, k)]
    K = int(rng.choice([1, 3, 8, 40, 96]))
    rmin = rmax * float(10 ** rng.uniform(-4, -0.3))
    _, r2 = t.oracle_thresholds(rmin, rmax, K)
    if K > 1 and rng.random() < 0.15:
        r2 = rng.permutation(r2).astype(np.float32)
    nfiles = int(rng.integers(1, 4))
    cuts = np.sort(rng.integers(0, npart, nfiles - 1))
    chunks = np.split(p, cuts)
    opts = {}
    pick = lambda key, vals: opts.__setitem__(key, int(rng.choice(vals)))
    if rng.random() < 0.6: pick("lanes_per_halo", [0, 8, 16, 64])
    if rng.random() < 0.5: pick("segment", [0, 4, 8])
    if rng.random() < 0.15: pick("kernel", [1])
    if rng.random() < 0.5: opts["cell_size_milli"] = int(rng.integers(3000, 30000))
    if rng.random() < 0.5: pick("x_split", [0, 1, 2, 3, 4, 8, 16])
    if rng.random() < 0.25: pick("index_sort", [1])
    if rng.random() < 0.4: pick("tiles", [0, 1])
    if rng.random() < 0.4: pick("tile_level", [0, 1, 2, 3, 4, 5, 6])
    if rng.random() < 0.3: pick("tile_halos", [0, 2, 3, 7, 64])
    if rng.random() < 0.3: pick("stage_particles", [256, 1024, 8192])
    if rng.random() < 0.3: pick("stage_halos", [0, 4, 16, 64, 1024])
    if rng.random() < 0.25: pick("graph", [1])
    if rng.random() < 0.25: pick("overlap", [1])
    if rng.random() < 0.2: opts["piece_max"] = int(rng.integers(1000, npart + 1))
    if rng.random() < 0.2: pick("fold_limit", [1, 5000])
    return chunks, h, r2, opts


@pytest.mark.parametrize("seed", range(160))
def test_random_settings(built, seed):
    chunks, h, r2, opts = draw_case(seed)
    print(f"case {seed}: particles {[c.shape[0] for c in chunks]} halos {h.shape[0]} "
          f"K {r2.size} r2max {float(r2.max()):.4g} options {opts}", flush=True)
    want = np.zeros((h.shape[0], r2.size), np.uint64)
    for c in chunks:
        want += t.oracle_count(c, h, r2)[0]
    e = built.Engine(0)
    try:
        for k, v in opts.items():
            e.set_option(k, v)
        e.set_shells(r2)
        e.set_halos(h)
        for c in chunks:
            e.add_particles(c)
        got = e.counts()
    finally:
        e.close()
    if not np.array_equal(got, want):
        rows = np.nonzero((got != want).any(axis=1))[0]
        raise AssertionError(f"{rows.size} halo rows differ, first {rows[:5]}: halo "
                             f"{h[rows[0]]} got {got[rows[0]]} want {want[rows[0]]}")

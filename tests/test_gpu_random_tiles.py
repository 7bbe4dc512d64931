"""Seeded random clustered workloads against the oracle: dense Gaussian
clumps, most halos exactly on clump centres (many per centre: tiles that
the cost pre-pass cuts into parts whose counts are added atomically), the
rest uniform or on faces, with a random choice of tile level, halos per
tile, lanes, segment, cell edge and x split.  The reference counts each
halo on its own (src/astr_rcount/main.c:795-826,
src/objects/ArraySTR/MultiRangeCount.c:13-115); tiles must not change a
count."""
import numpy as np
import pytest

import sjcs_testlib as t

pytestmark = pytest.mark.gpu


def draw_clustered(seed):
    rng = np.random.default_rng(60_000 + seed)
    npart = int(10 ** rng.uniform(4.0, 5.5))
    nc = int(rng.integers(2, 60))
    centres = (rng.random((nc, 3)) * 1000.0).astype(np.float32)
    sizes = rng.pareto(1.0, nc) + 1.0
    which = rng.choice(nc, npart, p=sizes / sizes.sum())
    p = centres[which] + rng.standard_normal((npart, 3)) * rng.uniform(0.2, 2.0)
    p = np.mod(p, 1000.0).astype(np.float32)
    p = np.minimum(p, np.nextafter(np.float32(1000.0), np.float32(0)))
    bg = int(npart * rng.uniform(0.0, 0.3))
    p[:bg] = rng.random((bg, 3), dtype=np.float32) * np.float32(1000.0)
    rmax = float(rng.uniform(1.0, 25.0))
    nh = int(min(10 ** rng.uniform(2.0, 3.8), 2e8 / npart))
    h = rng.random((nh, 3), dtype=np.float32) * np.float32(1000.0)
    on = rng.random(nh) < rng.uniform(0.5, 0.95)
    h[on] = centres[rng.integers(0, nc, int(on.sum()))]
    m = np.nonzero(rng.random(nh) < 0.05)[0]
    h[m, rng.integers(0, 3, m.size)] = rng.choice([0.0, 1000.0, 1.0, 999.0], m.size)
    K = int(rng.choice([1, 8, 40, 64]))
    _, r2 = t.oracle_thresholds(rmax * float(10 ** rng.uniform(-4, -0.5)), rmax, K)
    opts = {"tiles": 1}
    if rng.random() < 0.6: opts["tile_level"] = int(rng.choice([0, 1, 2, 3, 6]))
    if rng.random() < 0.4: opts["tile_halos"] = int(rng.choice([0, 2, 5, 16, 64]))
    if rng.random() < 0.5: opts["lanes_per_halo"] = int(rng.choice([0, 16, 64]))
    if rng.random() < 0.4: opts["segment"] = int(rng.choice([0, 4, 8]))
    if rng.random() < 0.4: opts["cell_size_milli"] = int(rng.integers(3000, 20000))
    if rng.random() < 0.4: opts["x_split"] = int(rng.choice([0, 1, 4, 8]))
    return p, h, r2, opts


@pytest.mark.parametrize("seed", range(40))
def test_random_clustered_tiles(built, seed):
    p, h, r2, opts = draw_clustered(seed)
    print(f"case {seed}: particles {p.shape[0]} halos {h.shape[0]} K {r2.size} "
          f"r2max {float(r2.max()):.4g} options {opts}", flush=True)
    want, _ = t.oracle_count(p, h, r2)
    e = built.Engine(0)
    try:
        for k, v in opts.items():
            e.set_option(k, v)
        e.set_shells(r2)
        e.set_halos(h)
        e.add_particles(p)
        got = e.counts()
    finally:
        e.close()
    if not np.array_equal(got, want):
        rows = np.nonzero((got != want).any(axis=1))[0]
        raise AssertionError(f"{rows.size} halo rows differ, first {rows[:5]}: halo "
                             f"{h[rows[0]]} got {got[rows[0]]} want {want[rows[0]]}")

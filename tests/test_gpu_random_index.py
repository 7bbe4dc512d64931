"""Seeded random index-API and serving workloads against the oracle.

Each case builds one index (include/objects/ArraySTR/ArraySTR.h:52-58:
construct, insert in one to three calls, make the tree) over uniform or
clumped particles on a fresh engine with a random choice of its options,
then counts several query batches -- each with its own centres (some on
faces, repeated, or on particles) and its own 1 to 96 thresholds up to
r_max 300 -- through sjcs_index_count (the per-halo periodic count,
src/objects/ArraySTR/MultiRangeCount.c:13-115) and through a serving loop
with one to eight slots, batches of 4096 halos or more included (the
loop's Morton-sorted path).  Every count must equal the oracle's."""
import numpy as np
import pytest

import sjcs_testlib as t

pytestmark = pytest.mark.gpu


def draw_points(rng, n):
    if rng.random() < 0.3:
        nc = int(rng.integers(3, 30))
        centres = rng.random((nc, 3)) * 1000.0
        p = centres[rng.integers(0, nc, n)] + rng.standard_normal((n, 3)) * rng.uniform(0.3, 3.0)
        p = np.mod(p, 1000.0).astype(np.float32)
        return np.minimum(p, np.nextafter(np.float32(1000.0), np.float32(0)))
    p = rng.random((n, 3), dtype=np.float32) * np.float32(1000.0)
    p[:4] = 1000.0
    p[4:8] = 0.0
    return p


def draw_batch(rng, p, npart):
    nh = int(rng.choice([1, 7, 300, 2000, 4096, 6000]))
    rmax = float(10 ** rng.uniform(-0.3, np.log10(300.0)))
    if nh * npart * min(1.0, (2 * rmax / 1000.0) ** 3) > 4e7 or (rmax > 60 and nh * npart > 2e7):
        nh = max(1, nh // 20)
    h = rng.random((nh, 3), dtype=np.float32) * np.float32(1000.0)
    m = np.nonzero(rng.random(nh) < 0.15)[0]
    h[m, rng.integers(0, 3, m.size)] = rng.choice([0.0, 1000.0, 2.5, 997.5], m.size)
    if nh > 10 and rng.random() < 0.4:
        h[rng.integers(0, nh, nh // 5)] = h[rng.integers(0, max(1, nh // 30), nh // 5)]
    if rng.random() < 0.3:
        k = max(1, nh // 10)
        h[rng.integers(0, nh, k)] = p[rng.integers(0, npart, k)]
    K = int(rng.choice([1, 3, 8, 40, 96]))
    _, r2 = t.oracle_thresholds(rmax * float(10 ** rng.uniform(-4, -0.3)), rmax, K)
    if K > 1 and rng.random() < 0.15:
        r2 = rng.permutation(r2).astype(np.float32)
    return h, r2


@pytest.mark.parametrize("seed", range(40))
def test_random_index_and_serving(built, seed):
    rng = np.random.default_rng(80_000 + seed)
    npart = int(10 ** rng.uniform(3.0, 5.0))
    p = draw_points(rng, npart)
    cuts = np.sort(rng.integers(0, npart, int(rng.integers(0, 3))))
    batches = [draw_batch(rng, p, npart) for _ in range(int(rng.integers(2, 6)))]
    opts = {}
    if rng.random() < 0.5: opts["lanes_per_halo"] = int(rng.choice([0, 8, 16, 64]))
    if rng.random() < 0.4: opts["segment"] = int(rng.choice([0, 4, 8]))
    if rng.random() < 0.4: opts["tiles"] = int(rng.choice([0, 1]))
    if rng.random() < 0.4: opts["tile_level"] = int(rng.choice([0, 1, 2, 4, 6]))
    if rng.random() < 0.3: opts["cell_size_milli"] = int(rng.integers(3000, 30000))
    if rng.random() < 0.3: opts["x_split"] = int(rng.choice([0, 1, 2, 4, 8]))
    slots = int(rng.integers(1, 9))
    print(f"case {seed}: particles {npart} inserts {cuts.size + 1} batches "
          f"{[(b[0].shape[0], b[1].size) for b in batches]} slots {slots} options {opts}",
          flush=True)
    want = [t.oracle_count(p, h, r2)[0] for h, r2 in batches]
    e = built.Engine(0)
    try:
        for k, v in opts.items():
            e.set_option(k, v)
        ix = built.ArraySTR(e, npart)
        for part in np.split(p, cuts):
            ix.insert(part)
        ix.make_tree()
        for (h, r2), w in zip(batches, want):
            got = ix.multi_range_count(h, r2)
            assert np.array_equal(got, w), f"index count differs: K {r2.size} halos {h.shape[0]}"
        sv = built.Server(ix, slots)
        try:
            tickets = [sv.submit(h, r2) for h, r2 in batches]
            for tk, (h, r2), w in zip(tickets, batches, want):
                assert np.array_equal(sv.wait(tk), w), \
                    f"served count differs: K {r2.size} halos {h.shape[0]}"
                sv.release(tk)
        finally:
            sv.close()
        ix.close()
    finally:
        e.close()

"""The clustered, uneven-work workload (200M particles, 70% in 200k Gaussian
clumps, x 5M halos, 80% at clump centres, --radius 0.01:20:64) against the
oracle: the autotuned setting, the same setting without halo tiles, and the
plain one-halo-per-wavefront kernel give the same counts, and those equal
the brute-force oracle -- on every halo of a reduced set drawn by the same
generator, and on a sample of the full-size set (the heaviest halos, clump
centres and field halos).

Reference semantics: src/objects/ArraySTR/MultiRangeCount.c:13-115 per halo,
the counts of every setting being the same set definition.

Runs on an MI355X only (marker `gpu`), through the C-ABI.
"""
import numpy as np
import pytest

import sjcs_testlib as t
from spatialjoincountovershells_amd import synth

pytestmark = pytest.mark.gpu


def assert_counts_equal(got, want, halos):
    if not np.array_equal(got, want):
        rows = np.nonzero((got != want).any(axis=1))[0]
        raise AssertionError(f"{rows.size} halo rows differ, first {rows[:5]}: "
                             f"halo {halos[rows[0]]} got {got[rows[0]]} want {want[rows[0]]}")


def three_settings(sj, p, h, r2, tag):
    """Counts of one chunk under the autotuned setting, the same setting
    with tiles off, and the plain kernel (default cells, no tiles); the
    tuned options and W are returned too."""
    e = sj.Engine(0)
    try:
        e.set_shells(r2)
        e.set_halos(h)
        d = e.device_array(p)
        try:
            ms = e.autotune(d, p.shape[0])
            opts = {k: e.get_option(k) for k in
                    ("cell_size_milli", "x_split", "lanes_per_halo", "segment", "tiles")}
            print(f"{tag}: autotuned {opts} ({ms:.2f} ms)", flush=True)
            e.reset_counts()
            e.add_particles_device(d, p.shape[0])
            tuned = e.counts()
            w = e.last_chunk_pair_tests()
            print(f"{tag}: tuned W {w}, pairs {int(tuned.sum())}", flush=True)
            e.set_option("tiles", 0)
            e.reset_counts()
            e.add_particles_device(d, p.shape[0])
            untiled = e.counts()
            assert e.last_chunk_pair_tests() == w
            print(f"{tag}: tiles off done", flush=True)
            for k in ("cell_size_milli", "x_split", "lanes_per_halo", "segment"):
                e.set_option(k, 0)
            e.set_option("kernel", 1)
            e.reset_counts()
            e.add_particles_device(d, p.shape[0])
            plain = e.counts()
            assert e.last_chunk_pair_tests() == w
            print(f"{tag}: plain kernel done", flush=True)
        finally:
            e.device_free(d)
    finally:
        e.close()
    return tuned, untiled, plain, opts, w


def test_clustered_reduced_every_halo(built):
    """2e6 particles (70% in 2000 clumps: as many members per clump as the
    full size has) x 2e4 halos, every halo against the oracle."""
    rng = np.random.default_rng(404)
    p, h = synth.clustered(rng, 2_000_000, 20_000, nclumps=2_000)
    _, r2 = t.oracle_thresholds(0.01, 20.0, 64)
    tuned, untiled, plain, opts, w = three_settings(built, p, h, r2, "reduced")
    want, w_oracle = t.oracle_count(p, h, r2, threads=16)
    print(f"reduced: oracle W {w_oracle}", flush=True)
    assert w == w_oracle
    assert_counts_equal(tuned, want, h)
    assert_counts_equal(untiled, want, h)
    assert_counts_equal(plain, want, h)
    assert int(want.sum()) > 0


def test_clustered_full_size(built):
    """The full workload: the three settings agree on all 5e6 x 64 counts,
    and 64 halos are checked against the oracle over all 2e8 particles: the
    16 heaviest distinct positions (most pairs counted: centres of the
    largest clumps, where tiles are cut into the most parts), 16 other
    clump-centre halos and 32 field halos."""
    rng = np.random.default_rng(2027)
    print("full clustered: generating", flush=True)
    p, h = synth.clustered(rng, 200_000_000, 5_000_000)
    _, r2 = t.oracle_thresholds(0.01, 20.0, 64)
    tuned, untiled, plain, opts, w = three_settings(built, p, h, r2, "full clustered")
    assert_counts_equal(untiled, tuned, h)
    assert_counts_equal(plain, tuned, h)
    del untiled, plain
    nc = int(round(0.8 * h.shape[0]))  # the generator puts the clump-centre halos first
    load = tuned.sum(axis=1)
    order = np.argsort(load[:nc])[::-1]
    _, first = np.unique(h[order], axis=0, return_index=True)
    heavy = order[np.sort(first)[:16]]
    others = np.setdiff1d(np.arange(nc), heavy)
    sample = np.concatenate([heavy, rng.choice(others, 16, replace=False),
                             nc + rng.choice(h.shape[0] - nc, 32, replace=False)])
    print(f"full clustered: oracle on {sample.size} halos, heaviest {load[heavy[0]]} pairs",
          flush=True)
    want, _ = t.oracle_count(p, h[sample], r2, threads=16)
    assert_counts_equal(tuned[sample], want, h[sample])

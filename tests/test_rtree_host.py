"""The library's host-side R-tree build (rt_rcount's index), CPU only: its
leaf grouping -- which decides rt_rcount's counts -- against the oracle's
restatement of src/objects/RTree/Build.c, itself pinned against rt_rcount's
results (test_oracle.py).  A host-only tree (no engine) needs no GPU."""
import zlib

import numpy as np
import pytest

import sjcs_testlib as t


def inputs(kind):
    rng = np.random.default_rng(zlib.crc32(kind.encode()))
    if kind == "uniform":
        return rng.random((40000, 3), dtype=np.float32) * np.float32(1000)
    if kind == "clustered":
        c = rng.random((50, 3)) * 1000
        p = c[rng.integers(0, 50, 30000)] + rng.normal(0, 1.5, (30000, 3))
        return np.mod(p, 1000).astype(np.float32)
    if kind == "grid":
        g = np.arange(28, dtype=np.float32)
        gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
        return np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1).astype(np.float32)
    if kind == "plane":  # zero volume everywhere: every area and enlargement ties
        p = rng.random((8000, 3), dtype=np.float32) * np.float32(1000)
        p[:, 2] = 7.0
        return p
    if kind == "duplicates":
        p = rng.random((2000, 3), dtype=np.float32) * np.float32(1000)
        return p[rng.integers(0, 2000, 20000)]
    if kind == "same_point":
        return np.full((3000, 3), 250.5, np.float32)
    if kind == "tie_file":  # grid points plus +-inf and +-1e30 coordinates
        return t.read_particles(t.GOLD + "/tie_particles.dat")
    if kind == "rt_file":
        return t.read_particles(t.GOLD + "/rt_particles.dat")
    if kind == "rtree_file":  # negative and out-of-box coordinates
        return t.read_particles(t.GOLD + "/rtree_particles.dat")
    raise ValueError(kind)


KINDS = ["uniform", "clustered", "grid", "plane", "duplicates", "same_point", "tie_file",
         "rt_file", "rtree_file"]


@pytest.mark.parametrize("kind", KINDS)
def test_rtree_leaves_match_oracle(built, kind):
    p = inputs(kind)
    with built.RTree() as rt:
        rt.insert(p)
        ids, lens = rt.leaves()
        shape = rt.shape()
    with t.OracleRTree(p) as o:
        oids, olens = o.leaves()
        assert shape["height"] == o.height()
    assert shape["points"] == p.shape[0] and shape["leaves"] == olens.size
    assert np.array_equal(lens, olens)
    assert np.array_equal(ids, oids)


def test_rtree_inserts_in_batches_equal_one_insert(built):
    """Ids continue across inserts and the points go in one at a time
    either way (rt_rcount/main.c:711-715)."""
    p = inputs("clustered")
    cuts = [0, 1, 10, 11, 500, 7777, 20000, p.shape[0]]
    with built.RTree() as rt:
        for a, b in zip(cuts[:-1], cuts[1:]):
            rt.insert(p[a:b])
        ids, lens = rt.leaves()
    with t.OracleRTree(p) as o:
        oids, olens = o.leaves()
    assert np.array_equal(ids, oids) and np.array_equal(lens, olens)


@pytest.mark.parametrize("n", [0, 1, 10, 11, 41])
def test_rtree_small_trees(built, n):
    """The root leaf, its first split (11 points), the first inner level."""
    p = inputs("uniform")[:n]
    with built.RTree() as rt:
        rt.insert(p)
        ids, lens = rt.leaves()
        shape = rt.shape()
    if n == 0:
        assert shape == {"points": 0, "leaves": 1, "nodes": 1, "height": 0}
        return
    with t.OracleRTree(p) as o:
        oids, olens = o.leaves()
        assert shape["height"] == o.height()
    assert np.array_equal(ids, oids) and np.array_equal(lens, olens)


def test_rtree_undefined_split_is_an_error(built):
    """All-NaN points: every enlargement difference is NaN, where the
    reference's PickNext returns no entry and DistributeEntry reads
    entries[-1] (src/objects/RTree/Build.c:77-101).  The library refuses."""
    p = np.full((11, 3), np.nan, np.float32)
    with built.RTree() as rt:
        rt.insert(p[:10])
        with pytest.raises(built.SJCSError) as ei:
            rt.insert(p[10:])
        assert "NaN" in str(ei.value)
        # the failed tree reports what went in before and refuses the rest
        assert rt.shape()["points"] == 10
        with pytest.raises(built.SJCSError) as ei:
            rt.leaves()
        assert ei.value.code == built.SJCS_ERR_STATE
    with pytest.raises(ValueError):
        t.OracleRTree(p)


def test_failed_insert_mid_batch(built):
    """A batch whose 11th point forces the undefined split: the points
    before it count, leaves() refuses (no buffer sized by the count could
    hold entries the split left behind)."""
    p = np.vstack([np.full((10, 3), np.nan, np.float32), np.full((5, 3), np.nan, np.float32)])
    with built.RTree() as rt:
        with pytest.raises(built.SJCSError):
            rt.insert(p)
        assert rt.shape()["points"] == 10
        with pytest.raises(built.SJCSError):
            rt.leaves()
        with pytest.raises(built.SJCSError):
            rt.insert(p[:1])


def test_host_only_tree_cannot_count(built):
    with built.RTree() as rt:
        rt.insert(inputs("uniform")[:100])
        with pytest.raises(built.SJCSError) as ei:
            rt.multi_range_count(np.zeros((1, 3), np.float32), np.ones(4, np.float32))
        assert ei.value.code == built.SJCS_ERR_STATE

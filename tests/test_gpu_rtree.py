"""rt_rcount's count on the GPU (include/sjcs.h, sjcs_rtree_*), bit-exact
against rt_rcount's own results (tests/golden/*.rt_expected.gz, made by the
reference binary built from its sources) and against the oracle's R-tree
restatement on larger inputs; and the drop-in CLI run as rt_rcount.

MI355X only (marker `gpu`); every call goes through the C-ABI or the CLI.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import sjcs_testlib as t

pytestmark = pytest.mark.gpu

RT_CASES = [c for c in t.manifest()["cases"] if "rt_expected" in c]


@pytest.fixture(scope="module")
def eng(built):
    e = built.Engine(0)
    yield e
    e.close()


def rt_count(built, eng, chunks, halos, r2):
    """One tree per particle file, counts summed over files (rt_rcount's
    chunk loop, src/rt_rcount/main.c:200-226)."""
    tot = np.zeros((halos.shape[0], len(r2)), np.uint64)
    for c in chunks:
        with built.RTree(eng) as rt:
            rt.insert(c)
            rt.multi_range_count(halos, r2, tot)
    return tot


def oracle_rt(chunks, halos, r2):
    tot = np.zeros((halos.shape[0], len(r2)), np.uint64)
    for c in chunks:
        tot += t.oracle_rt_count(c, halos, r2)
    return tot


def assert_counts_equal(got, want, halos):
    if not np.array_equal(got, want):
        rows = np.nonzero((got != want).any(axis=1))[0]
        raise AssertionError(f"{rows.size} halo rows differ, first {rows[:5]}: "
                             f"halo {halos[rows[0]]} got {got[rows[0]]} want {want[rows[0]]}")


@pytest.mark.parametrize("case", RT_CASES, ids=[c["name"] for c in RT_CASES])
def test_rt_api_golden_case(built, eng, case):
    print("case", case["name"], flush=True)
    mn, mx, k = t.parse_radius(case["radius"])
    _, r2 = t.oracle_thresholds(mn, mx, k)
    chunks, ids, pts = t.case_inputs(case)
    got = rt_count(built, eng, chunks, pts, r2)
    _, want, _, _ = t.parse_result(t.read_expected_text(case["rt_expected"]), k)
    assert_counts_equal(got, want, pts)


def on_particles(rng, p, n):
    return p[rng.integers(0, p.shape[0], n)]


def gen(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        p = rng.random((200_000, 3), dtype=np.float32) * np.float32(1000)
        h = np.vstack([on_particles(rng, p, 6000),
                       rng.random((14_000, 3), dtype=np.float32) * np.float32(1000)])
        h[-3000:, 0] = rng.uniform(0, 6, 3000).astype(np.float32)
        return p, h, (0.001, 5.0, 40)
    if kind == "clustered":
        c = rng.random((400, 3)) * 1000
        p = np.mod(c[rng.integers(0, 400, 120_000)] + rng.normal(0, 1.0, (120_000, 3)), 1000)
        p = np.vstack([p, rng.random((40_000, 3)) * 1000]).astype(np.float32)
        p = np.vstack([p, p[rng.integers(0, p.shape[0], 5000)]])  # duplicates
        h = np.vstack([c[:300], on_particles(rng, p, 4000),
                       rng.random((4000, 3)) * 1000]).astype(np.float32)
        return p, h, (0.01, 20.0, 64)
    if kind == "wide":  # query boxes wider than half the box: both bounds out, 8 images
        p = rng.random((30_000, 3), dtype=np.float32) * np.float32(1000)
        h = np.vstack([on_particles(rng, p, 50), rng.random((50, 3)) * 1000]).astype(np.float32)
        return p, h, (5.0, 600.0, 9)
    if kind == "odd_halos":  # outside the box, NaN, inf
        p = rng.random((20_000, 3), dtype=np.float32) * np.float32(1000)
        h = np.array([[-3, 500, 500], [1003, 2, 998], [np.nan, 5, 5], [np.inf, 5, 5],
                      [0, 0, 0], [1000, 1000, 1000], [-0.0, 999.9999, 0.0001]], np.float32)
        h = np.vstack([h, on_particles(rng, p, 100)])
        return p, h, (0.5, 30.0, 10)
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["uniform", "clustered", "wide", "odd_halos"])
def test_rt_api_matches_oracle(built, eng, kind):
    p, h, (mn, mx, k) = gen(kind, 41)
    print("rt", kind, p.shape[0], "points", h.shape[0], "halos", k, "shells", flush=True)
    _, r2 = t.oracle_thresholds(mn, mx, k)
    got = rt_count(built, eng, [p], h, r2)
    want = oracle_rt([p], h, r2)
    assert want.sum() > 0
    assert_counts_equal(got, want, h)


def test_rt_million_points(built, eng):
    """1e6 clustered and uniform points (height-6 tree), 2e4 halos, a
    quarter on particles and a tenth near faces, 40 shells, against the
    oracle's R-tree."""
    rng = np.random.default_rng(17)
    c = rng.random((2000, 3)) * 1000
    p = np.vstack([np.mod(c[rng.integers(0, 2000, 600_000)] + rng.normal(0, 3.0, (600_000, 3)),
                          1000), rng.random((400_000, 3)) * 1000]).astype(np.float32)
    h = np.vstack([on_particles(rng, p, 5000), rng.random((15_000, 3)) * 1000]).astype(np.float32)
    h[-2000:, 1] = rng.uniform(0, 4, 2000).astype(np.float32)
    _, r2 = t.oracle_thresholds(0.001, 5.0, 40)
    got = rt_count(built, eng, [p], h, r2)
    want = oracle_rt([p], h, r2)
    assert want.sum() > 1e5
    assert_counts_equal(got, want, h)


@pytest.mark.parametrize("shells", ["one", "unsorted", "many"])
def test_rt_shell_sets(built, eng, shells):
    """The first-match rule for any threshold order (linear scan) and the
    binary search for sorted ones, 1 to 1024 shells."""
    rng = np.random.default_rng(5)
    p = rng.random((50_000, 3), dtype=np.float32) * np.float32(1000)
    h = np.vstack([on_particles(rng, p, 500), rng.random((1500, 3)) * 1000]).astype(np.float32)
    if shells == "one":
        r2 = np.array([36.0], np.float32)
    elif shells == "unsorted":
        r2 = rng.permutation(t.oracle_thresholds(0.1, 12.0, 16)[1]).astype(np.float32)
    else:
        r2 = t.oracle_thresholds(0.01, 10.0, 1024)[1]
    print("shells", shells, r2.size, flush=True)
    got = rt_count(built, eng, [p], h, r2)
    assert_counts_equal(got, oracle_rt([p], h, r2), h)


@pytest.mark.parametrize("n", [0, 1, 10, 11, 150])
def test_rt_small_trees(built, eng, n):
    rng = np.random.default_rng(n)
    p = (rng.random((n, 3)) * 20 + 490).astype(np.float32)
    h = np.vstack([p[:3], rng.random((20, 3)) * 20 + 490]).astype(np.float32)
    _, r2 = t.oracle_thresholds(0.5, 8.0, 6)
    got = rt_count(built, eng, [p], h, r2)
    want = oracle_rt([p], h, r2) if n else np.zeros_like(got)
    assert_counts_equal(got, want, h)


def test_rt_insert_after_count(built, eng):
    """A count, more points, a count again: the device copy follows the
    tree, and the second count sees every point inserted so far."""
    rng = np.random.default_rng(9)
    p = rng.random((40_000, 3), dtype=np.float32) * np.float32(1000)
    h = np.vstack([on_particles(rng, p, 400), rng.random((600, 3)) * 1000]).astype(np.float32)
    _, r2 = t.oracle_thresholds(0.2, 9.0, 12)
    with built.RTree(eng) as rt:
        rt.insert(p[:15_000])
        first = rt.multi_range_count(h, r2)
        rt.insert(p[15_000:])
        second = rt.multi_range_count(h, r2)
        again = rt.multi_range_count(h, r2)
    assert_counts_equal(first, oracle_rt([p[:15_000]], h, r2), h)
    assert_counts_equal(second, oracle_rt([p], h, r2), h)
    assert np.array_equal(again, second)


def test_rt_device_entry_point_adds(built, eng):
    """sjcs_rtree_count_device: centres and counts on the device, counts
    added to what they hold (two calls: twice the host entry point's)."""
    rng = np.random.default_rng(12)
    p = rng.random((60_000, 3), dtype=np.float32) * np.float32(1000)
    h = np.vstack([on_particles(rng, p, 700), rng.random((1300, 3)) * 1000]).astype(np.float32)
    _, r2 = t.oracle_thresholds(0.05, 15.0, 24)
    with built.RTree(eng) as rt:
        rt.insert(p)
        want = rt.multi_range_count(h, r2)
        d_h = eng.device_array(h)
        d_c = eng.device_array(np.zeros((h.shape[0], r2.size), np.uint64))
        try:
            rt.multi_range_count_device(d_h, h.shape[0], r2, d_c)
            rt.multi_range_count_device(d_h, h.shape[0], r2, d_c)
            got = np.zeros_like(want)
            eng._check(eng._lib.sjcs_copy_to_host(eng._h, got.ctypes.data, d_c, got.nbytes))
        finally:
            eng.device_free(d_h)
            eng.device_free(d_c)
    assert_counts_equal(want, oracle_rt([p], h, r2), h)
    assert np.array_equal(got, 2 * want)


def test_rt_rejects_bad_shell_counts(built, eng):
    with built.RTree(eng) as rt:
        rt.insert(np.zeros((5, 3), np.float32))
        with pytest.raises(built.SJCSError):
            rt.multi_range_count(np.zeros((2, 3), np.float32), np.ones(1025, np.float32))


# ------------------------------------------------------- CLI as rt_rcount --

def mask(s):
    return re.sub(r"(>>>> [a-z +]+ time: )\d+( milli seconds)", r"\1#\2", s)


@pytest.fixture(scope="module")
def rt_cli(built, tmp_path_factory):
    """The drop-in CLI under the name rt_rcount (a symbolic link)."""
    d = tmp_path_factory.mktemp("rtbin")
    link = d / "rt_rcount"
    os.symlink(built.CLI_PATH, link)
    return str(link)


@pytest.mark.parametrize("case", RT_CASES, ids=[c["name"] for c in RT_CASES])
def test_rt_cli_golden_case(rt_cli, case, tmp_path):
    out = tmp_path / "result.txt"
    args = [rt_cli, "--thread", "4", "--next-setup", "2", "--particle-files"] + \
        case["particle_files"] + ["--halo-file", case["halo_file"], "--radius", case["radius"]]
    args += case["extra"] + ["--output-file", str(out)]
    print("case", case["name"], flush=True)
    r = subprocess.run(args, cwd=t.GOLD, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == t.read_expected_text(case["rt_expected"])
    assert mask(r.stdout) == case["rt_stdout_masked"]


def test_rt_cli_backlog(rt_cli, tmp_path):
    """--backlog: files still go one after the other, same result."""
    case = next(c for c in RT_CASES if c["name"] == "rt_2files")
    out = tmp_path / "result.txt"
    args = [rt_cli, "--thread", "4", "--next-setup", "1", "--backlog", "2", "--particle-files"] + \
        case["particle_files"] + ["--halo-file", case["halo_file"], "--radius", case["radius"],
                                  "--output-file", str(out)]
    r = subprocess.run(args, cwd=t.GOLD, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == t.read_expected_text(case["rt_expected"])
    assert "  -- with backlog execution" in r.stdout
    assert r.stdout.count(">>>> index getting time") == 2


@pytest.mark.parametrize("name", ["uni_8", "tdpb_4"])
@pytest.mark.parametrize("as_rt", [False, True])
def test_cli_result_on_stdout(built, rt_cli, name, as_rt):
    """No --output-file: outputResult writes the result to stdout
    (src/astr_rcount/main.c:383-385, rt_rcount/main.c the same), between
    the total searching time line and the result writing time line."""
    case = next(c for c in RT_CASES if c["name"] == name)
    exe = rt_cli if as_rt else built.CLI_PATH
    args = [exe, "--thread", "4", "--next-setup", "2", "--particle-files"] + \
        case["particle_files"] + ["--halo-file", case["halo_file"], "--radius", case["radius"]]
    r = subprocess.run(args + case["extra"], cwd=t.GOLD, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = case["rt_stdout_masked" if as_rt else "stdout_masked"]
    tail = ">>>> result writing time: # milli seconds\n"
    assert lines.endswith(tail)
    result = t.read_expected_text(case["rt_expected" if as_rt else "expected"]).decode()
    assert mask(r.stdout) == lines[:-len(tail)] + result + tail

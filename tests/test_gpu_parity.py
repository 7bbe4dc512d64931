"""Parity of the HIP path with the oracle and with the reference's golden
results.  Bit-exact: every count is an integer and must be equal.

Runs on an MI355X only (marker `gpu`); every call goes through the C-ABI
(lib/libsjcs.so, via ctypes) or the drop-in CLI that links it.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import sjcs_testlib as t

pytestmark = pytest.mark.gpu

CASES = t.manifest()["cases"]


@pytest.fixture(scope="module")
def eng(built):
    e = built.Engine(0)
    yield e
    e.close()


def gen_uniform(seed, n_particles, n_halos, face=0.1):
    """Config-1 style input: uniform particles (plus exact 0 and 1000
    coordinates), halos uniform with a fraction within 5 of a face."""
    rng = np.random.default_rng(seed)
    p = rng.random((n_particles, 3), dtype=np.float32) * np.float32(1000.0)
    p[:16] = 1000.0
    p[16:32] = 0.0
    h = rng.random((n_halos, 3), dtype=np.float32) * np.float32(1000.0)
    m = np.nonzero(rng.random(n_halos) < face)[0]
    ax = rng.integers(0, 3, m.size)
    lo = rng.random(m.size) < 0.5
    h[m, ax] = np.where(lo, rng.uniform(0, 5, m.size), rng.uniform(995, 1000, m.size))
    h[:3] = [[0, 0, 0], [1000, 1000, 1000], [999.99994, 0.0, 500.0]]
    return p, h


def run_engine(eng, chunks, halos, r2):
    eng.set_shells(r2)
    eng.set_halos(halos)
    for c in chunks:
        eng.add_particles(c)
    return eng.counts()


def oracle_chunks(chunks, halos, r2):
    tot = np.zeros((halos.shape[0], len(r2)), np.uint64)
    w = 0
    for c in chunks:
        cc, ww = t.oracle_count(c, halos, r2)
        tot += cc
        w += ww
    return tot, w


def face_weighted_sample(rng, h, n, rbox, frac=0.3):
    """n distinct halo indices, at least `frac` of them with a coordinate
    within rbox of a face (so their boxes leave the periodic box and they
    have images): every corner halo up to a tenth of those, edge halos (two
    axes near a face) up to a third, the rest face halos; the others drawn
    uniformly."""
    d = np.minimum(h, np.float32(1000.0) - h)
    nax = (d < rbox).sum(axis=1)
    k = int(np.ceil(frac * n))
    corner = np.nonzero(nax == 3)[0][: max(1, k // 10)]
    edge = np.nonzero(nax == 2)[0]
    edge = rng.choice(edge, min(edge.size, k // 3), replace=False)
    face = np.nonzero(nax == 1)[0]
    face = rng.choice(face, k - corner.size - edge.size, replace=False)
    near = np.concatenate([corner, edge, face])
    rest = np.setdiff1d(np.arange(h.shape[0]), near)
    rest = rng.choice(rest, n - near.size, replace=False)
    return np.concatenate([near, rest]), near.size


def assert_counts_equal(got, want, halos):
    if not np.array_equal(got, want):
        rows = np.nonzero((got != want).any(axis=1))[0]
        raise AssertionError(f"{rows.size} halo rows differ, first {rows[:5]}: "
                             f"halo {halos[rows[0]]} got {got[rows[0]]} want {want[rows[0]]}")


# ------------------------------------------------------------- golden CLI --

def mask(s):
    return re.sub(r"(>>>> [a-z +]+ time: )\d+( milli seconds)", r"\1#\2", s)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_cli_golden_case(built, case, tmp_path):
    out = tmp_path / "result.txt"
    args = [built.CLI_PATH, "--thread", "4", "--next-setup", "2", "--particle-files"] + \
        case["particle_files"] + ["--halo-file", case["halo_file"], "--radius", case["radius"]]
    if case.get("filter_file"):
        args += ["--filter-file", case["filter_file"]]
    args += case["extra"] + ["--output-file", str(out)]
    print("case", case["name"], flush=True)
    r = subprocess.run(args, cwd=t.GOLD, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    if not case.get("filter_file"):
        assert out.read_bytes() == t.read_expected_text(case["expected"])
        assert mask(r.stdout) == case["stdout_masked"]
        return
    # --filter-file: the reference reads an uninitialised id count
    # (src/Gadget/BufHalo.c:251-300 never stores it).  Its golden file is
    # what that count being 0 gives; this engine keeps every listed halo.
    mn, mx, k = t.parse_radius(case["radius"])
    _, r2 = t.oracle_thresholds(mn, mx, k)
    chunks, ids, pts = t.case_inputs(case)
    want, _ = oracle_chunks(chunks, pts, r2)
    gids, gcnt, gtot, _ = t.parse_result(out.read_bytes(), k)
    assert np.array_equal(gids, ids) and np.array_equal(gcnt, want)
    _, rids, _ = t.case_inputs(case, filter_as_reference_runs=True)
    eids, *_ = t.parse_result(t.read_expected_text(case["expected"]), k)
    assert np.array_equal(eids, rids)


def write_halo_list(path, ids, pts):
    """The reference's halo text (src/Gadget/BufHalo.c:76-119): 16 header
    lines, then id and 7 more columns with x, y, z in columns 9-11; %.9g
    gives back the same FP32 values through strtof."""
    with open(path, "w") as f:
        f.writelines(f"# header {i}\n" for i in range(16))
        f.writelines(f"{i} 1.5 0 0 0 0 0 0 {x:.9g} {y:.9g} {z:.9g} 0\n"
                     for i, (x, y, z) in zip(ids.tolist(), pts.tolist()))


@pytest.mark.parametrize("threads", [1, 16])
def test_cli_result_file_many_batches(built, tmp_path, threads):
    """The drop-in CLI end to end on 140000 halos (five batches of the
    result writer, the counts fetched from the device batch by batch while
    the previous one is formatted) and two particle files: its result file
    is byte for byte the plain formatter's text of the oracle's counts."""
    p, h = gen_uniform(51, 400_000, 140_000)
    ids = np.arange(140_000, dtype=np.uint64) * np.uint64(7919) + np.uint64(3)
    files = [tmp_path / "p0.dat", tmp_path / "p1.dat"]
    p[:250_000].tofile(files[0])
    p[250_000:].tofile(files[1])
    hf = tmp_path / "halos.list"
    write_halo_list(hf, ids, h)
    out = tmp_path / "result.txt"
    args = [built.CLI_PATH, "--thread", str(threads), "--next-setup", "1", "--particle-files",
            str(files[0]), str(files[1]), "--halo-file", str(hf), "--radius", "0.001:5:8",
            "--output-file", str(out)]
    print("cli many batches, threads", threads, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _, r2 = t.oracle_thresholds(0.001, 5.0, 8)
    want, _ = oracle_chunks([p[:250_000], p[250_000:]], h, r2)
    got = out.read_bytes()
    expect = t.format_result(ids, want, r2)
    assert len(got) == len(expect) and got == expect


@pytest.mark.parametrize("backlog", [["--backlog", "1"], ["--backlog", "2", "--first-setup", "2:1"],
                                     ["--backlog", "5", "--first-setup", "1:3"]])
def test_cli_backlog_modes(built, backlog, tmp_path):
    """--backlog N holds up to N particle files on the device ahead of the
    one being counted: the result file is still the reference's, and stdout
    has the backlog mode's lines (main.c:292-323): one "index n" per file,
    an "index getting time" and a "searching time" per file."""
    case = [c for c in CASES if c["name"] == "uni_3files"][0]
    out = tmp_path / "result.txt"
    args = [built.CLI_PATH, "--thread", "4", "--next-setup", "1", "--particle-files"] + \
        case["particle_files"] + ["--halo-file", case["halo_file"], "--radius", case["radius"]] + \
        backlog + ["--output-file", str(out)]
    r = subprocess.run(args, cwd=t.GOLD, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == t.read_expected_text(case["expected"])
    n = len(case["particle_files"])
    assert "  -- with backlog execution" in r.stdout
    assert [ln for ln in r.stdout.splitlines() if ln.startswith("index ")] == \
        [f"index {i}" for i in range(n)]
    assert r.stdout.count(">>>> index getting time") == n
    assert r.stdout.count(">>>> searching time") == n
    assert r.stdout.count(">>>> particles reading time") == n


def test_dpps_equivalence_on_the_gpu_host():
    """The DPPS probe (the FP32 distance as the reference's SSE4.1 dot
    product rounds it, against the oracle's plain expression) on the host
    CPU of the machine with the GPU: the CPU whose reference binary the
    benchmark times."""
    t.oracle()
    out = subprocess.run([t.DPPS_PROBE, "3000000", "4242"], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("dpps ok")


@pytest.mark.parametrize("missing", ["particles", "halos", "second_particles", "compressed"])
def test_cli_missing_inputs_with_device(built, missing, tmp_path):
    """With the device set up on its own thread, a missing input still ends
    the run with the reference's message and status 1 (no crash in the
    runtime's teardown)."""
    parts = ["uni_particles.dat"]
    halo = "uni_halos.list"
    extra = []
    if missing == "particles":
        parts = ["nope.dat"]
    elif missing == "halos":
        halo = "nope.list"
    elif missing == "second_particles":
        parts = ["uni_particles.dat", "nope.dat"]
    else:
        parts = ["nope.gdt"]
        extra = ["--compressed"]
    args = [built.CLI_PATH, "--thread", "4", "--next-setup", "2", "--particle-files"] + parts + \
        extra + ["--halo-file", halo, "--radius", "0.5:5:4", "--output-file", str(tmp_path / "r")]
    r = subprocess.run(args, cwd=t.GOLD, capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, (r.returncode, r.stderr)
    want = {"particles": "ERROR: file nope.dat not found in readDecodedParticles()",
            "halos": "ERROR: file nope.list not found in readHalos().",
            "second_particles": "ERROR: file nope.dat not found in readDecodedParticles()",
            "compressed": "ERROR: file nope.gdt not found in readCompressedParticles()"}[missing]
    assert want in r.stderr


# ---------------------------------------------------------- engine parity --

@pytest.mark.parametrize("seed", [1, 2])
def test_config1_exact(eng, seed):
    """Config 1 at its own size: 1e5 particles, 1e4 halos (10% near faces),
    --radius 0.001:5:8."""
    p, h = gen_uniform(seed, 100_000, 10_000)
    _, r2 = t.oracle_thresholds(0.001, 5.0, 8)
    got = run_engine(eng, [p], h, r2)
    want, w = oracle_chunks([p], h, r2)
    assert_counts_equal(got, want, h)
    assert eng.last_chunk_pair_tests() == w


@pytest.mark.parametrize("spec", ["0.001:5:40", "0.5:100:16", "0.01:20:64", "1:600:5",
                                  "3:7:1", "0.25:2:7", "20:0.01:64"])
def test_shell_sets(eng, spec):
    p, h = gen_uniform(11, 20_000, 2_000)
    mn, mx, k = t.parse_radius(spec)
    _, r2 = t.oracle_thresholds(mn, mx, k)
    got = run_engine(eng, [p], h, r2)
    want, w = oracle_chunks([p], h, r2)
    assert_counts_equal(got, want, h)
    assert eng.last_chunk_pair_tests() == w


def test_many_shells_count_in_hbm(eng):
    """K above the LDS histogram limit: counts go straight to HBM atomics."""
    p, h = gen_uniform(12, 8_000, 400)
    _, r2 = t.oracle_thresholds(0.01, 40.0, 1500)
    got = run_engine(eng, [p], h, r2)
    assert_counts_equal(got, oracle_chunks([p], h, r2)[0], h)


@pytest.mark.parametrize("kind", ["nan_inside", "inf_top", "descending", "zeros", "neg_zero",
                                  "nan_top", "duplicates"])
def test_unusual_shell_thresholds(eng, kind):
    """Thresholds the CLI never makes but the ABI accepts: the first-match
    rule must still hold exactly (general path, no lookup table)."""
    p, h = gen_uniform(13, 10_000, 800)
    _, r2 = t.oracle_thresholds(0.01, 8.0, 10)
    r2 = r2.copy()
    if kind == "nan_inside":
        r2[3] = np.nan
    elif kind == "inf_top":
        r2[-1] = np.inf
        r2[-2] = np.inf
    elif kind == "descending":
        r2 = r2[::-1].copy()
    elif kind == "zeros":
        r2[:3] = 0.0
    elif kind == "neg_zero":
        r2[0] = -0.0
    elif kind == "nan_top":
        r2[-1] = np.nan
        p = p[:2000]
        h = h[:100]
    elif kind == "duplicates":
        r2[4:7] = r2[5]
    got = run_engine(eng, [p], h, r2)
    assert_counts_equal(got, oracle_chunks([p], h, r2)[0], h)


def test_special_coordinates(eng):
    """Halos and particles at box faces, outside the box, huge, infinite and
    NaN (a NaN halo counts nothing; NaN particles are left out because the
    reference's tree gives layout-dependent results for them)."""
    rng = np.random.default_rng(14)
    p = rng.random((30_000, 3), dtype=np.float32) * np.float32(1000.0)
    specials = np.array([[0, 0, 0], [1000, 1000, 1000], [-2, 500, 500], [1002, 3, 4],
                         [1e30, 1, 1], [-1e30, 2, 2], [np.inf, 3, 3], [-np.inf, 999, 1],
                         [0, 1000, 0], [999.99994, 0, 0], [-0.0, 5, 5]], np.float32)
    p[:specials.shape[0]] = specials
    h = rng.random((1_000, 3), dtype=np.float32) * np.float32(1000.0)
    h[:10] = [[np.nan, 1, 1], [np.inf, 5, 5], [-np.inf, 5, 5], [1e30, 1, 1], [0, 0, 0],
              [1000, 1000, 1000], [-3, 500, 500], [1003, 2, 2], [500, np.nan, 500],
              [2.5, 997.5, 0.0]]
    _, r2 = t.oracle_thresholds(0.001, 6.0, 12)
    got = run_engine(eng, [p], h, r2)
    want, w = oracle_chunks([p], h, r2)
    assert_counts_equal(got, want, h)
    assert got[0].sum() == 0 and got[8].sum() == 0
    # a NaN centre's box passes every particle on its NaN axis: W counts them
    assert eng.last_chunk_pair_tests() == w


@pytest.mark.parametrize("index_sort", [0, 1])
@pytest.mark.parametrize("cell_milli", [977, 1500, 2500, 3907, 4999, 10000, 50000, 250000,
                                        1000000])
def test_cell_size_does_not_change_counts(eng, cell_milli, index_sort):
    """Every cell size, with the per-axis sort (up to 256 cells per axis)
    and with the packed-key sort."""
    p, h = gen_uniform(15, 30_000, 2_000)
    _, r2 = t.oracle_thresholds(0.001, 5.0, 8)
    want, _ = oracle_chunks([p], h, r2)
    eng.set_option("cell_size_milli", cell_milli)
    eng.set_option("index_sort", index_sort)
    try:
        got = run_engine(eng, [p], h, r2)
    finally:
        eng.set_option("cell_size_milli", 0)
        eng.set_option("index_sort", 0)
    assert_counts_equal(got, want, h)


@pytest.mark.parametrize("index_sort", [0, 1])
@pytest.mark.parametrize("cell_milli,x_split", [(4500, 1), (4500, 2), (4500, 5), (4500, 16),
                                                (2000, 7), (977, 3), (30000, 16)])
def test_x_split_does_not_change_counts(eng, cell_milli, x_split, index_sort):
    """Finer cells along x than along y/z (up to 1024 of them), with both
    sorts."""
    p, h = gen_uniform(22, 30_000, 2_000, face=0.3)
    _, r2 = t.oracle_thresholds(0.001, 5.0, 12)
    want, _ = oracle_chunks([p], h, r2)
    eng.set_option("cell_size_milli", cell_milli)
    eng.set_option("x_split", x_split)
    eng.set_option("index_sort", index_sort)
    try:
        got = run_engine(eng, [p], h, r2)
    finally:
        eng.set_option("cell_size_milli", 0)
        eng.set_option("x_split", 0)
        eng.set_option("index_sort", 0)
    assert_counts_equal(got, want, h)


@pytest.mark.parametrize("lanes,segment", [(8, 4), (8, 8), (16, 4), (16, 8), (64, 4), (64, 8)])
def test_kernel_shapes(eng, lanes, segment):
    """Every lanes-per-halo / segment shape of the count kernel, on halos near
    faces and corners (periodic images), with a one-step and a multi-step
    shell table."""
    p, h = gen_uniform(21, 40_000, 3_000, face=0.3)
    eng.set_option("lanes_per_halo", lanes)
    eng.set_option("segment", segment)
    try:
        for spec in ("0.001:5:40", "0.01:20:64", "1:600:5"):
            mn, mx, k = t.parse_radius(spec)
            _, r2 = t.oracle_thresholds(mn, mx, k)
            got = run_engine(eng, [p], h, r2)
            want, w = oracle_chunks([p], h, r2)
            assert_counts_equal(got, want, h)
            assert eng.last_chunk_pair_tests() == w
    finally:
        eng.set_option("lanes_per_halo", 0)
        eng.set_option("segment", 0)


@pytest.mark.parametrize("lanes,segment,stage,halos", [
    (16, 4, 2048, 0), (16, 8, 4096, 16), (8, 4, 1024, 32), (16, 4, 256, 64), (8, 8, 8192, 128)])
def test_particle_staging(eng, lanes, segment, stage, halos):
    """Particles staged in LDS per workgroup (the union of its halos' base
    boxes): the counts equal the oracle's, with images across the periodic
    boundary read from HBM, workgroups whose union does not fit (small
    stage_particles) falling back, dense clumps and sparse regions."""
    p, h = gen_uniform(60 + lanes + segment, 150_000, 6_000)
    rng = np.random.default_rng(61)
    clump = (np.float32(400.0) + rng.normal(0, 2.0, (40_000, 3))).astype(np.float32)
    p = np.concatenate([p, clump]).astype(np.float32)
    h[:300] = (np.float32(400.0) + rng.uniform(-3, 3, (300, 3))).astype(np.float32)
    _, r2 = t.oracle_thresholds(0.01, 6.0, 24)
    want, _ = oracle_chunks([p], h, r2)
    try:
        eng.set_option("lanes_per_halo", lanes)
        eng.set_option("segment", segment)
        eng.set_option("stage_particles", stage)
        eng.set_option("stage_halos", halos)
        got = run_engine(eng, [p], h, r2)
    finally:
        for k in ("lanes_per_halo", "segment", "stage_particles", "stage_halos"):
            eng.set_option(k, 0)
    assert_counts_equal(got, want, h)


@pytest.mark.parametrize("index_sort", [0, 1])
def test_clustered_particles(eng, index_sort):
    """Most particles in one tight clump (whole sort tiles of one cell, one
    digit per wavefront), halos around and inside it, plus a uniform
    background; chunk sizes off the tile size."""
    rng = np.random.default_rng(19)
    clump = (np.float32(500.0) + rng.normal(0, 0.4, (70_001, 3))).astype(np.float32)
    bg = rng.random((9_999, 3), dtype=np.float32) * np.float32(1000.0)
    p = np.concatenate([clump[:40_000], bg, clump[40_000:]])
    h = (np.float32(500.0) + rng.uniform(-6, 6, (300, 3))).astype(np.float32)
    h = np.concatenate([h, rng.random((200, 3), dtype=np.float32) * np.float32(1000.0)])
    _, r2 = t.oracle_thresholds(0.001, 5.0, 8)
    eng.set_option("index_sort", index_sort)
    try:
        got = run_engine(eng, [p[:16_385], p[16_385:]], h, r2)
    finally:
        eng.set_option("index_sort", 0)
    want, _ = oracle_chunks([p[:16_385], p[16_385:]], h, r2)
    assert_counts_equal(got, want, h)


def long_rows_input(seed, lengths, clumped):
    """Uniform background plus rows of the 64 x 64 (y, z) cell grid (15.625
    wide) holding exactly `lengths` particles; rows in `clumped` put all
    their added particles in one x cell (500.1 .. 500.4)."""
    G = 64
    cs = 1000.0 / G
    inv = np.float32(G) / np.float32(1000.0)
    rng = np.random.default_rng(seed)

    def row_of(q):
        c = np.clip(np.floor(q[:, 1:] * inv), 0, G - 1).astype(np.int64)
        return c[:, 1] * G + c[:, 0]

    bg = rng.random((20_000, 3), dtype=np.float32) * np.float32(1000.0)
    rows = rng.choice(G * G, len(lengths), replace=False)
    parts, halos = [bg], [rng.random((100, 3), dtype=np.float32) * np.float32(1000.0)]
    have = np.bincount(row_of(bg), minlength=G * G)
    for i, (rw, ln) in enumerate(zip(rows, lengths)):
        cz, cy = divmod(int(rw), G)
        m = ln - int(have[rw])
        q = np.empty((m, 3), np.float32)
        q[:, 0] = (rng.uniform(500.1, 500.4, m) if i in clumped
                   else rng.uniform(0.0, 1000.0, m))
        q[:, 1] = rng.uniform(cy * cs + 0.5, (cy + 1) * cs - 0.5, m)
        q[:, 2] = rng.uniform(cz * cs + 0.5, (cz + 1) * cs - 0.5, m)
        parts.append(q)
        hq = np.empty((40, 3), np.float32)
        hq[:, 0] = rng.uniform(490.0, 510.0, 40) if i in clumped else rng.uniform(0, 1000, 40)
        hq[:, 1] = rng.uniform(cy * cs, (cy + 1) * cs, 40)
        hq[:, 2] = rng.uniform(cz * cs, (cz + 1) * cs, 40)
        halos.append(hq)
    p = np.concatenate(parts)
    p = p[rng.permutation(p.shape[0])]
    got_len = np.bincount(row_of(p), minlength=G * G)[rows]
    assert got_len.tolist() == list(lengths)
    return p, np.concatenate(halos)


@pytest.mark.parametrize("x_split", [16, 3, 1])
def test_long_rows(eng, x_split):
    """Rows of the last sort pass just under, at and over the LDS row size
    (4096 particles) and over several 8192-particle tiles, along x uniform or
    all in one x cell, with 1024, 192 and 64 cells along x: rows over 4096
    particles take the tiled long-row path."""
    lengths = [4095, 4096, 4097, 8192, 8193, 3 * 8192 + 5, 30_001, 12_000]
    p, h = long_rows_input(31, lengths, clumped={7})
    _, r2 = t.oracle_thresholds(0.001, 5.0, 8)
    want, _ = oracle_chunks([p], h, r2)
    eng.set_option("cell_size_milli", 15625)
    eng.set_option("x_split", x_split)
    try:
        got = run_engine(eng, [p], h, r2)
    finally:
        eng.set_option("cell_size_milli", 0)
        eng.set_option("x_split", 0)
    assert_counts_equal(got, want, h)
    assert want.sum() > 0


def test_chunk_accumulation_and_reset(eng):
    """Counts add over chunks (empty and one-particle chunks included) and
    restart from zero after reset."""
    p, h = gen_uniform(16, 50_000, 3_000)
    _, r2 = t.oracle_thresholds(0.001, 5.0, 8)
    chunks = [p[:0], p[:1], p[1:20_000], p[20_000:20_001], p[20_001:]]
    got = run_engine(eng, chunks, h, r2)
    assert_counts_equal(got, oracle_chunks(chunks, h, r2)[0], h)
    eng.reset_counts()
    assert eng.counts().sum() == 0
    eng.add_particles(p[:10_000])
    assert_counts_equal(eng.counts(), oracle_chunks([p[:10_000]], h, r2)[0], h)


def test_pipelined_chunks(eng):
    """Chunks alternate between host and device input and grow in size (the
    two index slots, staging buffers and workspace reallocate mid-stream);
    a reset, a new halo set and new shells between chunk streams."""
    p, h = gen_uniform(23, 120_000, 3_000, face=0.2)
    _, r2 = t.oracle_thresholds(0.001, 5.0, 16)
    cuts = [0, 5_000, 5_001, 25_000, 60_000, 61_000, 120_000]
    chunks = [p[a:b] for a, b in zip(cuts[:-1], cuts[1:])]
    eng.set_shells(r2)
    eng.set_halos(h)
    dev = [eng.device_array(c) if i % 2 else None for i, c in enumerate(chunks)]
    try:
        for c, d in zip(chunks, dev):
            if d is None:
                eng.add_particles(c)
            else:
                eng.add_particles_device(d, c.shape[0])
        assert_counts_equal(eng.counts(), oracle_chunks(chunks, h, r2)[0], h)
        eng.reset_counts()
        for c, d in zip(chunks[::-1], dev[::-1]):
            if d is None:
                eng.add_particles(c)
            else:
                eng.add_particles_device(d, c.shape[0])
        h2 = h[::2].copy()
        _, r2b = t.oracle_thresholds(0.01, 8.0, 9)
        eng.set_halos(h2)
        eng.set_shells(r2b)
        for c in chunks[:3]:
            eng.add_particles(c)
        assert_counts_equal(eng.counts(), oracle_chunks(chunks[:3], h2, r2b)[0], h2)
    finally:
        for d in dev:
            if d is not None:
                eng.device_free(d)


@pytest.mark.parametrize("limit", [1, 45_000])
def test_count_folds(eng, limit):
    """Device counts run in u32 and fold into u64 totals before they could
    overflow; a tiny fold limit folds after every chunk (or every other), and
    read-out adds the u64 totals and the running u32 counts."""
    p, h = gen_uniform(24, 90_000, 2_500, face=0.2)
    _, r2 = t.oracle_thresholds(0.001, 5.0, 10)
    chunks = [p[:20_000], p[20_000:40_000], p[40_000:41_000], p[41_000:]]
    eng.set_option("fold_limit", limit)
    try:
        got = run_engine(eng, chunks, h, r2)
        assert_counts_equal(got, oracle_chunks(chunks, h, r2)[0], h)
        assert np.array_equal(eng.counts(), got)  # read-out repeats
        eng.reset_counts()
        eng.add_particles(chunks[0])
        assert_counts_equal(eng.counts(), oracle_chunks(chunks[:1], h, r2)[0], h)
    finally:
        eng.set_option("fold_limit", 0)


def test_graph_chunk_loop(eng):
    """Graph mode: each chunk's sort + count is captured once and replayed.
    Every replay (device and host input, store and accumulate graphs) equals
    the eager run, with eager kernels and synchronisations between replays;
    a new halo set recaptures."""
    p, h = gen_uniform(25, 70_000, 2_500, face=0.2)
    _, r2 = t.oracle_thresholds(0.001, 5.0, 12)
    chunks = [p[:30_000], p[30_000:]]
    eager = run_engine(eng, chunks, h, r2)
    assert_counts_equal(eager, oracle_chunks(chunks, h, r2)[0], h)
    w_eager = eng.last_chunk_pair_tests()
    d = [eng.device_array(c) for c in chunks]
    eng.set_option("graph", 1)
    try:
        for rep in range(4):
            eng.reset_counts()
            for c, dd in zip(chunks, d):
                eng.add_particles_device(dd, c.shape[0])
            assert np.array_equal(eng.counts(), eager), f"device replay {rep}"
            assert eng.last_chunk_pair_tests() == w_eager
            eng.reset_counts()
            for c in chunks:
                eng.add_particles(c)
            assert np.array_equal(eng.counts(), eager), f"host replay {rep}"
        h2 = h[::3].copy()
        eng.set_halos(h2)
        eng.reset_counts()
        for c, dd in zip(chunks, d):
            eng.add_particles_device(dd, c.shape[0])
        assert_counts_equal(eng.counts(), oracle_chunks(chunks, h2, r2)[0], h2)
    finally:
        eng.set_option("graph", 0)
        for dd in d:
            eng.device_free(dd)


def tiled_halo_set(rng):
    """Halos that form tiles (runs of >= 4 sharing a Morton cell of the
    1024^3 grid): exact duplicates at a box corner (periodic images), a run of
    70 spread inside one cell (split at the tile cap), NaN centres (Morton
    cell 0) next to real halos of cell 0, and uniform singles."""
    m = np.float32(1000.0 / 1024.0)
    corner = np.repeat(np.array([[0.3, 999.8, 0.2]], np.float32), 40, axis=0)
    cell = ((np.array([300, 611, 47], np.float32) + np.float32(0.1)) * m +
            rng.uniform(0, 0.8, (70, 3)).astype(np.float32) * m)
    nanrun = np.array([[np.nan, 1, 1]] * 5 + [[0.1, 0.2, 0.3]] * 3, np.float32)
    singles = rng.random((500, 3), dtype=np.float32) * np.float32(1000.0)
    return np.concatenate([singles[:200], corner, cell, nanrun, singles[200:]]).astype(np.float32)


@pytest.mark.parametrize("spec", ["0.001:5:40", "0.01:20:64", "0.5:100:16", "1:600:5"])
def test_halo_tiles(eng, spec):
    """Tiles (shared candidates, one wavefront per tile) against the oracle,
    and equal to the one-halo-at-a-time count, W included."""
    rng = np.random.default_rng(26)
    p = rng.random((60_000, 3), dtype=np.float32) * np.float32(1000.0)
    clump = (np.array([1.0, 999.0, 2.0], np.float32) +
             rng.normal(0, 1.5, (20_000, 3))).astype(np.float32) % np.float32(1000.0)
    p = np.concatenate([p, clump]).astype(np.float32)
    h = tiled_halo_set(rng)
    mn, mx, k = t.parse_radius(spec)
    _, r2 = t.oracle_thresholds(mn, mx, k)
    got = run_engine(eng, [p], h, r2)
    w_tiles = eng.last_chunk_pair_tests()
    want, w = oracle_chunks([p], h, r2)
    assert_counts_equal(got, want, h)
    assert w_tiles == w
    eng.set_option("tiles", 0)
    try:
        assert np.array_equal(run_engine(eng, [p], h, r2), got)
        assert eng.last_chunk_pair_tests() == w
    finally:
        eng.set_option("tiles", 1)


@pytest.mark.parametrize("level,halos", [(1, 0), (2, 0), (3, 8), (4, 2), (6, 0), (3, 64)])
def test_tile_levels(eng, level, halos):
    """Tiles of neighbouring halos (halos in one Morton cell of 2^level grid
    cells share their candidates), cut to at most `halos` per tile: counts
    and W equal the oracle's, on clustered particles with halos at clump
    centres, in the field and at the box corner."""
    rng = np.random.default_rng(30 + level)
    p = rng.random((60_000, 3), dtype=np.float32) * np.float32(1000.0)
    clump = (np.array([1.0, 999.0, 2.0], np.float32) +
             rng.normal(0, 1.5, (20_000, 3))).astype(np.float32) % np.float32(1000.0)
    p = np.concatenate([p, clump]).astype(np.float32)
    h = tiled_halo_set(rng)
    near = (np.array([500.0, 500.0, 500.0], np.float32) +
            rng.uniform(-12, 12, (300, 3))).astype(np.float32)
    h = np.concatenate([h, near]).astype(np.float32)
    _, r2 = t.oracle_thresholds(0.01, 20.0, 64)
    want, w = oracle_chunks([p], h, r2)
    try:
        eng.set_option("tile_level", level)
        eng.set_option("tile_halos", halos)
        got = run_engine(eng, [p], h, r2)
        assert_counts_equal(got, want, h)
        assert eng.last_chunk_pair_tests() == w
    finally:
        eng.set_option("tile_level", 0)
        eng.set_option("tile_halos", 0)


def test_tile_level_changes_keep_counts(eng):
    """Count rows do not depend on how halos are tiled: the tile level may
    change between particle files while counts accumulate."""
    rng = np.random.default_rng(36)
    p, h = gen_uniform(37, 120_000, 4_000)
    h[:400] = (np.float32(300.0) + rng.uniform(-6, 6, (400, 3))).astype(np.float32)
    _, r2 = t.oracle_thresholds(0.01, 8.0, 16)
    eng.set_shells(r2)
    eng.set_halos(h)
    try:
        for i, level in enumerate((0, 3, 1, 4)):
            eng.set_option("tile_level", level)
            eng.add_particles(p[i * 30_000:(i + 1) * 30_000])
        got = eng.counts()
    finally:
        eng.set_option("tile_level", 0)
    assert_counts_equal(got, oracle_chunks([p], h, r2)[0], h)


def test_autotune_cache(built):
    """A second autotune of the same workload takes the first one's setting
    from the cache without timing; another workload (other shells) is
    tuned afresh; with the cache off every call tunes."""
    rng = np.random.default_rng(38)
    p = rng.random((60_000, 3), dtype=np.float32) * np.float32(1000.0)
    h = np.concatenate([np.repeat(np.array([[200.0, 300.0, 400.0]], np.float32), 9, axis=0),
                        rng.random((800, 3), dtype=np.float32) * np.float32(1000.0)])
    keys = ("cell_size_milli", "x_split", "lanes_per_halo", "segment", "tiles", "tile_level",
            "tile_halos")
    e = built.Engine(0)
    try:
        _, r2 = t.oracle_thresholds(0.02, 9.0, 24)
        e.set_shells(r2)
        e.set_halos(h)
        d = e.device_array(p)
        try:
            ms1 = e.autotune(d, p.shape[0])
            assert e.get_option("autotune_cache_hit") == 0
            first = {k: e.get_option(k) for k in keys}
            for k in ("cell_size_milli", "x_split", "tile_level"):
                e.set_option(k, 0)
            ms2 = e.autotune(d, p.shape[0])
            assert e.get_option("autotune_cache_hit") == 1
            assert ms2 == ms1
            assert {k: e.get_option(k) for k in keys} == first
            e.add_particles_device(d, p.shape[0])
            assert_counts_equal(e.counts(), oracle_chunks([p], h, r2)[0], h)
            _, r2b = t.oracle_thresholds(0.02, 9.5, 24)
            e.set_shells(r2b)
            e.autotune(d, p.shape[0])
            assert e.get_option("autotune_cache_hit") == 0
            e.set_option("autotune_cache", 0)
            e.set_shells(r2)
            e.autotune(d, p.shape[0])
            assert e.get_option("autotune_cache_hit") == 0
        finally:
            e.device_free(d)
    finally:
        e.close()


def test_autotune_keeps_counts(built):
    """The autotuner picks settings from timed runs; counts are unchanged by
    it and equal the oracle afterwards, on clustered data with tiles."""
    rng = np.random.default_rng(27)
    p = rng.random((80_000, 3), dtype=np.float32) * np.float32(1000.0)
    clump = (np.float32(300.0) + rng.normal(0, 1.0, (30_000, 3))).astype(np.float32)
    p = np.concatenate([p, clump]).astype(np.float32)
    h = np.concatenate([np.repeat(np.array([[300.0, 300.0, 300.0]], np.float32), 12, axis=0),
                        rng.random((1_500, 3), dtype=np.float32) * np.float32(1000.0)])
    _, r2 = t.oracle_thresholds(0.01, 20.0, 16)
    e = built.Engine(0)
    try:
        e.set_shells(r2)
        e.set_halos(h)
        d = e.device_array(p)
        try:
            e.add_particles_device(d, p.shape[0])
            before = e.counts()
            ms = e.autotune(d, p.shape[0])
            assert ms > 0
            assert np.array_equal(e.counts(), before)  # counts untouched
            for key in ("cell_size_milli", "x_split", "lanes_per_halo", "segment", "tiles"):
                e.get_option(key)
            e.reset_counts()
            e.add_particles_device(d, p.shape[0])
            assert_counts_equal(e.counts(), oracle_chunks([p], h, r2)[0], h)
        finally:
            e.device_free(d)
    finally:
        e.close()


def test_server_alone_equals_mix(built, eng):
    """Serving loop: requests of 1 to 30000 halos, each with its own shells
    (K 8-64, r_max 1-10, r_min = r_max/5000), give the same counts alone
    (submitted and awaited one by one) and in the mix (all submitted at once,
    more than slots and budget hold), equal to the oracle on the small ones."""
    rng = np.random.default_rng(28)
    p = rng.random((300_000, 3), dtype=np.float32) * np.float32(1000.0)
    ix = built.ArraySTR(eng, p.shape[0])
    ix.insert(p)
    ix.make_tree()
    reqs = []
    for n in (1, 3, 64, 500, 4095, 4096, 9000, 30_000, 2, 700, 12_000, 1):
        rmax = float(rng.uniform(1.0, 10.0))
        k = int(rng.integers(8, 65))
        _, r2 = t.oracle_thresholds(rmax / 5000.0, rmax, k)
        h = rng.random((n, 3), dtype=np.float32) * np.float32(1000.0)
        reqs.append((h, r2))
    sv = built.Server(ix, slots=3, mem_budget=64 << 20)
    try:
        alone = []
        for h, r2 in reqs:
            tk = sv.submit(h, r2)
            alone.append(sv.wait(tk).copy())
            sv.release(tk)
        tickets = [sv.submit(h, r2) for h, r2 in reqs]
        for (h, r2), a, tk in zip(reqs, alone, tickets):
            assert np.array_equal(sv.wait(tk), a)
            sub, adm, done = sv.times(tk)
            assert sub <= adm <= done
        for (h, r2), a in zip(reqs, alone):
            if h.shape[0] <= 700:
                want, _ = t.oracle_count(p, h, r2, threads=8)
                assert_counts_equal(a, want, h)
            else:
                # requests of 4096 halos or more are Morton-ordered and read
                # out through the permutation: equal to the index API (whose
                # own test checks it against the oracle), and a sample of
                # rows against the oracle directly
                assert_counts_equal(a, ix.multi_range_count(h, r2), h)
                rows = np.random.default_rng(h.shape[0]).choice(h.shape[0], 300, replace=False)
                want, _ = t.oracle_count(p, h[rows], r2, threads=8)
                assert_counts_equal(a[rows], want, h[rows])
    finally:
        sv.close()
        ix.close()


def test_server_concurrent_submitters(built, eng):
    """Four host threads submit and await requests on one server at the same
    time (counts copied out in pieces by the server's copy threads while
    others are launched): every result equals the same request served
    alone."""
    import threading
    rng = np.random.default_rng(29)
    p = rng.random((200_000, 3), dtype=np.float32) * np.float32(1000.0)
    ix = built.ArraySTR(eng, p.shape[0])
    ix.insert(p)
    ix.make_tree()
    reqs = []
    for i in range(48):
        n = int(10 ** rng.uniform(0.0, 4.7))
        rmax = float(rng.uniform(1.0, 10.0))
        k = int(rng.integers(8, 65))
        _, r2 = t.oracle_thresholds(rmax / 5000.0, rmax, k)
        reqs.append((rng.random((n, 3), dtype=np.float32) * np.float32(1000.0), r2))
    sv = built.Server(ix, slots=4)
    try:
        alone = []
        for h, r2 in reqs:
            tk = sv.submit(h, r2)
            alone.append(sv.wait(tk).copy())
            sv.release(tk)
        got = [None] * len(reqs)
        errors = []

        def worker(w):
            try:
                mine = list(range(w, len(reqs), 4))
                tks = [(i, sv.submit(*reqs[i])) for i in mine]
                for i, tk in tks:
                    got[i] = sv.wait(tk).copy()
                    sv.release(tk)
            except Exception as exc:  # reported below, on the main thread
                errors.append(exc)

        threads = [threading.Thread(target=worker, args=(w,)) for w in range(4)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        assert not errors, errors
        for i, a in enumerate(alone):
            assert np.array_equal(got[i], a), i
    finally:
        sv.close()
        ix.close()


def test_degenerate_sizes(eng):
    _, r2 = t.oracle_thresholds(0.1, 5.0, 4)
    p, h = gen_uniform(17, 5_000, 50)
    assert run_engine(eng, [p], h[:0], r2).shape == (0, 4)
    assert run_engine(eng, [p[:0]], h, r2).sum() == 0
    one = np.array([[500.0, 500.0, 500.0]], np.float32)
    got = run_engine(eng, [one], one, r2)
    assert got.tolist() == [[1, 0, 0, 0]]


def test_index_api_matches_oracle(built, eng):
    """The ArraySTR-shaped lifecycle: construct, insert in pieces, build,
    count two shell sets into caller-owned counts (+=)."""
    p, h = gen_uniform(18, 40_000, 1_500)
    ix = built.ArraySTR(eng, p.shape[0])
    ix.insert(p[:10_000])
    ix.insert(p[10_000:])
    ix.make_tree()
    for spec in ("0.001:5:8", "0.5:30:12"):
        mn, mx, k = t.parse_radius(spec)
        _, r2 = t.oracle_thresholds(mn, mx, k)
        counts = np.full((h.shape[0], k), 7, np.uint64)
        ix.multi_range_count(h, r2, counts)
        want, _ = oracle_chunks([p], h, r2)
        assert_counts_equal(counts, want + np.uint64(7), h)
    ix.close()


@pytest.mark.parametrize("level", [0, 2])
def test_index_api_tiles(built, eng, level):
    """The index count tiles its query halos like the chunk loop: halos at
    one spot (runs of copies) and, at a tile level above 0, near one another
    share candidates; counts equal the oracle and the untiled count."""
    rng = np.random.default_rng(40 + level)
    p = rng.random((100_000, 3), dtype=np.float32) * np.float32(1000.0)
    clump = (np.float32(700.0) + rng.normal(0, 1.0, (30_000, 3))).astype(np.float32)
    p = np.concatenate([p, clump]).astype(np.float32)
    h = np.concatenate([np.repeat(np.array([[700.0, 700.0, 700.0]], np.float32), 30, axis=0),
                        (np.float32(700.0) + rng.uniform(-4, 4, (40, 3))).astype(np.float32),
                        tiled_halo_set(rng)]).astype(np.float32)
    h = h[rng.permutation(h.shape[0])]
    _, r2 = t.oracle_thresholds(0.01, 12.0, 32)
    ix = built.ArraySTR(eng, p.shape[0])
    try:
        ix.insert(p)
        ix.make_tree()
        eng.set_option("tile_level", level)
        got = ix.multi_range_count(h, r2)
        eng.set_option("tiles", 0)
        plain = ix.multi_range_count(h, r2)
    finally:
        eng.set_option("tiles", 1)
        eng.set_option("tile_level", 0)
        ix.close()
    want, _ = oracle_chunks([p], h, r2)
    assert_counts_equal(got, want, h)
    assert np.array_equal(plain, got)


# ------------------------------------------------ full size, by properties --

# ----------------------------------------------- compressed (Gadget) input --

def gadget_params(built, total, box, dmin, dmax, bits, id_start, ntile=1, fused=1):
    from test_oracle import gadget_header
    prm = built.gadget_decode_params(gadget_header(total, box, dmin, dmax, bits, id_start),
                                     ntile)
    prm.fused = fused
    return prm


def as_dict(prm):
    return {f: getattr(prm, f) for f in ("id_start", "ng1d", "nt1d", "cell_interval",
                                         "disp_min", "unitsep")}


DECODE_SETS = {
    # whole grid and ids past it (a second tile along x)
    "grid24": (24 ** 3, 1000.0, -30.0, 30.0, 16, 1, 1, 2_000_000),
    # 2048^3 (total in npartTotal[2]), ids across the grid's end
    "big2048": (2048 ** 3, 1000.0, -1.5, 2.5, 14, 2 ** 33 - 500_000, 1, 1_000_000),
    # the largest side, ids just below 2^50
    "side46340": (46340 ** 3, 1000.0, -0.25, 0.25, 16, 2 ** 50 - 1_500_000, 1, 1_000_000),
    # ids below 1: C's truncating division of negative numbers
    "negative": (100 ** 3, 1000.0, -5.0, 5.0, 16, -700_000, 1, 1_500_000),
    # three tiles per side
    "tiles3": (300 ** 3, 900.0, -3.0, 1.0, 12, 1, 3, 3_000_000),
    # spacing inexact in FP32, fused and unfused differ
    "fma48": (48 ** 3, 1000.0, -7.3, 7.3, 16, 1, 1, 110_592),
}


@pytest.mark.parametrize("fused", [1, 0])
@pytest.mark.parametrize("name", sorted(DECODE_SETS))
def test_decode_bit_exact(built, eng, name, fused):
    total, box, dmin, dmax, bits, id0, ntile, n = DECODE_SETS[name]
    prm = gadget_params(built, total, box, dmin, dmax, bits, id0, ntile, fused)
    rng = np.random.default_rng(len(name) + fused)
    uv = rng.integers(0, 1 << bits, (n, 3), dtype=np.uint16)
    uv[:7] = [0, 0, 0]
    uv[7:14] = (1 << bits) - 1
    print("decode", name, "fused", fused, "n", n, flush=True)
    got = eng.decode_compressed(uv, prm)
    want = t.oracle_gadget_decode(uv, as_dict(prm), fused)
    bad = np.nonzero((got.view(np.uint32) != want.view(np.uint32)).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} particles differ, first {bad[:4]}: {got[bad[0]]} {want[bad[0]]}"
    if name == "fma48":
        other = t.oracle_gadget_decode(uv, as_dict(prm), 1 - fused)
        assert not np.array_equal(got, other)


def test_decode_degenerate_sizes(built, eng):
    prm = gadget_params(built, 24 ** 3, 1000.0, -30.0, 30.0, 16, 5)
    assert eng.decode_compressed(np.zeros((0, 3), np.uint16), prm).shape == (0, 3)
    uv = np.array([[1, 2, 65535]], np.uint16)
    assert np.array_equal(eng.decode_compressed(uv, prm), t.oracle_gadget_decode(uv, as_dict(prm)))


def test_decode_rejects_what_it_cannot_decode_exactly(built, eng):
    uv = np.zeros((4, 3), np.uint16)
    cases = [("ng1d", 0), ("ng1d", 46341), ("nt1d", 0), ("nt1d", 46341),
             ("id_start", 2 ** 50), ("id_start", 1 - 2 ** 50)]
    for field, value in cases:
        prm = gadget_params(built, 24 ** 3, 1000.0, -1.0, 1.0, 16, 1)
        setattr(prm, field, value)
        with pytest.raises(built.SJCSError) as ei:
            eng.decode_compressed(uv, prm)
        assert ei.value.code == built.SJCS_ERR_INVALID, (field, value)
    # grid points past 2^31: side 10, ids near 3e11 (tile number 3e8)
    prm = gadget_params(built, 10 ** 3, 1000.0, -1.0, 1.0, 16, 3 * 10 ** 11)
    with pytest.raises(built.SJCSError) as ei:
        eng.decode_compressed(uv, prm)
    assert ei.value.code == built.SJCS_ERR_INVALID
    # the engine stays usable, and a chunk it refuses adds nothing
    _, r2 = t.oracle_thresholds(0.5, 30.0, 10)
    eng.set_shells(r2)
    eng.set_halos(np.array([[500.0, 500.0, 500.0]], np.float32))
    with pytest.raises(built.SJCSError):
        eng.add_compressed_particles(uv, prm)
    assert int(eng.counts().sum()) == 0


@pytest.mark.parametrize("graph", [0, 1])
def test_compressed_files_equal_oracle(built, eng, graph):
    """The golden two-file snapshot through the engine's compressed path:
    counts equal the oracle's on its own decode."""
    _, pts = t.read_halos(os.path.join(t.GOLD, "uni_halos.list"))
    _, r2 = t.oracle_thresholds(0.5, 30.0, 10)
    files = [os.path.join(t.GOLD, f) for f in ("gad_part0.gdt", "gad_part1.gdt")]
    eng.set_option("graph", graph)
    try:
        eng.set_shells(r2)
        eng.set_halos(pts)
        for f in files:
            hdr, uv = built.read_compressed(f)
            eng.add_compressed_particles(uv, built.gadget_decode_params(hdr))
        got = eng.counts()
    finally:
        eng.set_option("graph", 0)
    want, _ = oracle_chunks([t.oracle_decode_file(f) for f in files], pts, r2)
    assert_counts_equal(got, want, pts)


def test_compressed_chunks_equal_decoded_chunks(built, eng):
    """Chunks of growing and shrinking size through both slots: the
    compressed path counts exactly what the decoded path counts for the
    oracle's decode of the same triplets."""
    rng = np.random.default_rng(77)
    ng = 110
    _, r2 = t.oracle_thresholds(0.001, 5.0, 8)
    h = rng.random((20_000, 3), dtype=np.float32) * np.float32(1000.0)
    chunks = []
    start = 1
    for n in (400_000, 1_100_000, 250_000, 600_000):
        prm = gadget_params(built, ng ** 3, 1000.0, -4.0, 4.0, 16, start)
        chunks.append((rng.integers(0, 65536, (n, 3), dtype=np.uint16), prm))
        start += n
    eng.set_shells(r2)
    eng.set_halos(h)
    for uv, prm in chunks:
        eng.add_compressed_particles(uv, prm)
    got = eng.counts()
    eng.reset_counts()
    for uv, prm in chunks:
        eng.add_particles(t.oracle_gadget_decode(uv, as_dict(prm)))
    want = eng.counts()
    assert_counts_equal(got, want, h)
    assert int(got.sum()) > 0


SORT_DECODE_SETS = dict(DECODE_SETS, **{
    # one tile, ids mid-grid, partial sort tiles (the cell sort's own decode)
    "grid200": (200 ** 3, 1000.0, -4.0, 4.0, 16, 5_000_001, 1, 301_111),
})


@pytest.mark.parametrize("fused", [1, 0])
@pytest.mark.parametrize("name", sorted(SORT_DECODE_SETS))
def test_compressed_sort_equals_decoded(built, eng, name, fused):
    """Files of one tile per side are decoded by the cell sort's first pass
    as it reads them, the others by the decode kernel before the sort: both
    count what the oracle's decode of the same triplets counts, with halos
    next to particles so that the coordinates' low bits matter."""
    total, box, dmin, dmax, bits, id0, ntile, n = SORT_DECODE_SETS[name]
    n = min(n, 400_000)
    prm = gadget_params(built, total, box, dmin, dmax, bits, id0, ntile, fused)
    rng = np.random.default_rng(900 + 2 * len(name) + fused)
    uv = rng.integers(0, 1 << bits, (n, 3), dtype=np.uint16)
    p = t.oracle_gadget_decode(uv, as_dict(prm), fused)
    # halos where the particles are, also outside the box (ids below 1 put
    # the particles at negative coordinates)
    h = (p[rng.integers(0, n, 4000)] + rng.normal(0.0, 0.05, (4000, 3))).astype(np.float32)
    _, r2 = t.oracle_thresholds(0.001, 2.0, 16)
    print("sort decode", name, "fused", fused, "n", n, flush=True)
    eng.set_shells(r2)
    eng.set_halos(h)
    eng.add_compressed_particles(uv, prm)
    got = eng.counts()
    eng.reset_counts()
    eng.add_particles(p)
    want = eng.counts()
    assert_counts_equal(got, want, h)
    assert int(got.sum()) > 0


def test_compressed_device_chunks(built, eng):
    """Triplets already on the device give what host triplets give."""
    rng = np.random.default_rng(78)
    _, r2 = t.oracle_thresholds(0.01, 8.0, 16)
    h = rng.random((5_000, 3), dtype=np.float32) * np.float32(1000.0)
    hdr, uv = synth_snapshot(rng, 60)
    prm = built.gadget_decode_params(hdr)
    eng.set_shells(r2)
    eng.set_halos(h)
    eng.add_compressed_particles(uv, prm)
    want = eng.counts()
    d = eng.device_array(uv)
    try:
        eng.reset_counts()
        eng.add_compressed_particles_device(d, uv.shape[0], prm)
        eng.add_compressed_particles_device(d, uv.shape[0] // 3, prm)
        got = eng.counts()
        eng.reset_counts()
        eng.add_compressed_particles_device(d, uv.shape[0] // 3, prm)
        third = eng.counts()
    finally:
        eng.device_free(d)
    assert_counts_equal(got, want + third, h)
    assert int(want.sum()) > 0


@pytest.mark.parametrize("graph", [0, 1])
def test_pieces_equal_whole_chunks(built, eng, graph):
    """A chunk cut into pieces (piece_max, an odd size, so pieces end inside
    sort tiles) counts what it counts whole: packed particles, a compressed
    file of one tile per side (each piece decoded by its sort from its own
    first id) and one of three tiles per side (decoded before the sort)."""
    rng = np.random.default_rng(4242)
    _, r2 = t.oracle_thresholds(0.001, 4.0, 12)
    p, h = gen_uniform(43, 250_000, 3_000)
    q, _ = gen_uniform(44, 260_000, 3)
    h = np.vstack([h, p[:1500] + np.float32(0.5)]).astype(np.float32)
    one = gadget_params(built, 100 ** 3, 1000.0, -4.0, 4.0, 16, 1)
    many = gadget_params(built, 300 ** 3, 900.0, -3.0, 1.0, 12, 1, 3)
    uv1 = rng.integers(0, 1 << 16, (300_000, 3), dtype=np.uint16)
    uv3 = rng.integers(0, 1 << 12, (300_000, 3), dtype=np.uint16)
    # with 70001 particles per piece: 250k, 260k and 280k particles are 4
    # pieces each -- an even number, so the next chunk is uploaded into the
    # staging that the previous chunk's last piece (run from the other slot)
    # may still be reading; 300k are 5 pieces
    uv1b = rng.integers(0, 1 << 16, (280_000, 3), dtype=np.uint16)
    uv3b = rng.integers(0, 1 << 12, (280_000, 3), dtype=np.uint16)
    eng.set_shells(r2)
    eng.set_halos(h)
    res = []
    try:
        eng.set_option("graph", graph)
        for cap in (0, 70_001):
            print("pieces graph", graph, "piece_max", cap, flush=True)
            eng.set_option("piece_max", cap)
            assert eng.get_option("piece_max") == cap
            eng.reset_counts()
            eng.add_particles(p)
            eng.add_particles(q)
            eng.add_compressed_particles(uv1b, one)
            eng.add_compressed_particles(uv1, one)
            eng.add_compressed_particles(uv3b, many)
            eng.add_compressed_particles(uv3b[::-1], many)
            eng.add_compressed_particles(uv3, many)
            eng.add_particles(q[::-1])
            res.append(eng.counts())
    finally:
        eng.set_option("piece_max", 0)
        eng.set_option("graph", 0)
    assert_counts_equal(res[1], res[0], h)
    assert int(res[0].sum()) > 0


def synth_snapshot(rng, side):
    from spatialjoincountovershells_amd import synth
    return synth.gadget_snapshot(rng, side)


def test_counts_range_follows_the_totals(eng):
    """Row ranges equal the full read-out, and a range read after more
    chunks (or a reset) sees the new totals, not a stale read-out."""
    p, h = gen_uniform(31, 60_000, 5_000)
    _, r2 = t.oracle_thresholds(0.01, 10.0, 12)
    eng.set_shells(r2)
    eng.set_halos(h)
    eng.add_particles(p[:30_000])
    full = eng.counts()
    parts = [eng.counts_range(a, b - a) for a, b in ((0, 1), (1, 2500), (2500, 4999), (4999, 5000))]
    assert np.array_equal(np.vstack(parts), full)
    assert eng.counts_range(5000, 0).shape == (0, 12)
    eng.add_particles(p[30_000:])
    want, _ = oracle_chunks([p], h, r2)
    assert_counts_equal(eng.counts_range(0, 5000), want, h)
    eng.reset_counts()
    assert int(eng.counts_range(100, 50).sum()) == 0


def test_counts_range_u32(built, eng):
    """The 32-bit rows equal the 64-bit ones while the totals are the
    device's u32 running counts (one chunk, then a second one without a
    fold), in any range and with the widths alternating; after a fold into
    64 bits they are refused (SJCS_ERR_STATE) until a reset, after which they
    are zeros and then the new chunk's counts."""
    p, h = gen_uniform(37, 60_000, 5_000, face=0.2)
    _, r2 = t.oracle_thresholds(0.01, 10.0, 12)
    eng.set_shells(r2)
    eng.set_halos(h)
    assert eng.counts_fit_u32()
    assert int(eng.counts_range_u32(0, 5000).sum()) == 0
    eng.add_particles(p[:30_000])
    eng.add_particles(p[30_000:])
    assert eng.counts_fit_u32()
    want, _ = oracle_chunks([p], h, r2)
    got = eng.counts_range_u32(0, 5000)
    assert got.dtype == np.uint32
    assert_counts_equal(got.astype(np.uint64), want, h)
    for a, b in ((0, 1), (1, 2500), (2500, 4999), (4999, 5000)):
        assert np.array_equal(eng.counts_range(a, b - a), want[a:b])
        assert np.array_equal(eng.counts_range_u32(a, b - a).astype(np.uint64), want[a:b])
    assert eng.counts_range_u32(5000, 0).shape == (0, 12)
    eng.set_option("fold_limit", 1)
    try:
        eng.reset_counts()
        eng.add_particles(p[:30_000])
        eng.add_particles(p[30_000:])  # folds the first chunk into 64 bits
        assert not eng.counts_fit_u32()
        with pytest.raises(built.SJCSError) as err:
            eng.counts_range_u32(0, 10)
        assert err.value.code == built.SJCS_ERR_STATE
        assert_counts_equal(eng.counts_range(0, 5000), want, h)
        eng.reset_counts()
        assert eng.counts_fit_u32()
        assert int(eng.counts_range_u32(0, 5000).sum()) == 0
        eng.add_particles(p[:30_000])
        one, _ = oracle_chunks([p[:30_000]], h, r2)
        assert_counts_equal(eng.counts_range_u32(0, 5000).astype(np.uint64), one, h)
    finally:
        eng.set_option("fold_limit", 0)


def test_full_size_config2_properties(eng):
    """Config 2 at its full size (1e8 particles x 1e7 halos x 40 shells),
    where the oracle is too slow: counts are independent of the cell size
    and of how the particles are split into chunks, equal the oracle on a
    sample of halos, and W is near its expectation 1e9."""
    rng = np.random.default_rng(2026)
    p = rng.random((100_000_000, 3), dtype=np.float32) * np.float32(1000.0)
    h = rng.random((10_000_000, 3), dtype=np.float32) * np.float32(1000.0)
    _, r2 = t.oracle_thresholds(0.001, 5.0, 40)
    d = eng.device_array(p)
    try:
        eng.set_shells(r2)
        eng.set_halos(h)
        eng.add_particles_device(d, p.shape[0])
        base = eng.counts()
        # ranges above the 64 MiB download piece, at an odd offset and at
        # the end, equal the same rows of the whole result
        for a, m in ((1_234_567, 300_001), (h.shape[0] - 250_000, 250_000)):
            assert np.array_equal(eng.counts_range(a, m), base[a:a + m])
        # the 32-bit rows the CLI reads back at this size: the same values
        assert eng.counts_fit_u32()
        for a, m in ((0, 32_768), (1_234_567, 300_001), (h.shape[0] - 250_000, 250_000)):
            assert np.array_equal(eng.counts_range_u32(a, m).astype(np.uint64), base[a:a + m])
        w = eng.last_chunk_pair_tests()
        assert abs(w / 1e9 - 1.0) < 0.005
        print("full size: W", w, "pairs", int(base.sum()), flush=True)
        eng.set_option("cell_size_milli", 5000)
        eng.reset_counts()
        eng.add_particles_device(d, p.shape[0])
        assert np.array_equal(eng.counts(), base)
        eng.set_option("index_sort", 1)
        eng.reset_counts()
        eng.add_particles_device(d, p.shape[0])
        assert np.array_equal(eng.counts(), base)
        eng.set_option("index_sort", 0)
        eng.set_option("cell_size_milli", 0)
        eng.reset_counts()
        third = p.shape[0] // 3
        for a, b in ((0, third), (third, 2 * third), (2 * third, p.shape[0])):
            eng.add_particles_device(d + 12 * a, b - a)
        assert np.array_equal(eng.counts(), base)
    finally:
        eng.set_option("cell_size_milli", 0)
        eng.set_option("index_sort", 0)
        eng.device_free(d)
    sample, nnear = face_weighted_sample(rng, h, 1000, 5.0)
    print(f"full size: oracle on {sample.size} halos ({nnear} within r_box of a face)",
          flush=True)
    want, _ = t.oracle_count(p, h[sample], r2, threads=16)
    assert_counts_equal(base[sample], want, h[sample])


def test_full_size_config3_properties(eng):
    """The largest uniform workload: 3 particle files of 1e8 uniform particles
    (distinct seeds) x 1e7 halos x 40 shells, counts accumulated on the
    device.  The 3-file totals equal the sum of each file counted alone,
    equal the graph-replayed chunk loop, equal the oracle on a sample of
    halos, and W is near its expectation 3e9."""
    rng = np.random.default_rng(3003)
    h = rng.random((10_000_000, 3), dtype=np.float32) * np.float32(1000.0)
    _, r2 = t.oracle_thresholds(0.001, 5.0, 40)
    files = [np.random.default_rng(3100 + i).random((100_000_000, 3), dtype=np.float32) *
             np.float32(1000.0) for i in range(3)]
    ds = [eng.device_array(p) for p in files]
    try:
        eng.set_shells(r2)
        eng.set_halos(h)
        w = 0
        each = []
        for d in ds:
            eng.reset_counts()
            eng.add_particles_device(d, 100_000_000)
            each.append(eng.counts())
            w += eng.last_chunk_pair_tests()
        print("config 3: W", w, flush=True)
        assert abs(w / 3e9 - 1.0) < 0.005
        eng.reset_counts()
        for d in ds:
            eng.add_particles_device(d, 100_000_000)
        total = eng.counts()
        assert np.array_equal(total, each[0] + each[1] + each[2])
        eng.set_option("graph", 1)
        eng.reset_counts()
        for d in ds:
            eng.add_particles_device(d, 100_000_000)
        assert np.array_equal(eng.counts(), total)
    finally:
        eng.set_option("graph", 0)
        for d in ds:
            eng.device_free(d)
    sample, nnear = face_weighted_sample(rng, h, 500, 5.0)
    print(f"config 3: oracle on {sample.size} halos ({nnear} within r_box of a face)",
          flush=True)
    want = np.zeros((sample.size, 40), np.uint64)
    for i, p in enumerate(files):
        want += t.oracle_count(p, h[sample], r2, threads=16)[0]
        print(f"config 3: oracle file {i} done", flush=True)
    assert_counts_equal(total[sample], want, h[sample])

"""Pins the oracle (oracle/sjcs_oracle.c), CPU only.

Against: the golden vectors the compiled reference produced
(tests/golden/), the hand-derived periodic known answer of the reference's
TestData4PB, the SSE4.1 DPPS instruction itself, and closed forms on
integer lattices -- each chosen so that a plausible slip (a dropped periodic
image, a shifted sign, <= for <, an FMA in the distance, a swapped axis)
fails at least one of them.
"""
import itertools
import subprocess

import numpy as np
import pytest

import sjcs_testlib as t

CASES = t.manifest()["cases"]


def oracle_case_counts(case, as_reference_runs=True):
    mn, mx, k = t.parse_radius(case["radius"])
    _, r2 = t.oracle_thresholds(mn, mx, k)
    chunks, ids, pts = t.case_inputs(case, filter_as_reference_runs=as_reference_runs)
    tot = np.zeros((pts.shape[0], k), np.uint64)
    for ch in chunks:
        c, _ = t.oracle_count(ch, pts, r2)
        tot += c
    return ids, tot, k


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_equals_reference_golden(case):
    ids, tot, k = oracle_case_counts(case)
    eids, ecnt, etot, _ = t.parse_result(t.read_expected_text(case["expected"]), k)
    assert np.array_equal(eids, ids)
    assert np.array_equal(ecnt, tot)
    # Total = sum over halos + halo 0 once more (main.c:400,413)
    want_total = tot.sum(axis=0, dtype=np.uint64) + (tot[0] if tot.shape[0] else 0)
    assert np.array_equal(etot, want_total)


def test_tdpb_known_answer():
    """TestData4PB with --radius 0.001:5:8, derived by hand: each halo sees
    the 8 corner particles through its images at d^2 = 0, 1 (x3), 2 (x3), 3."""
    case = [c for c in CASES if c["name"] == "tdpb_8"][0]
    ids, tot, k = oracle_case_counts(case)
    per_halo = np.array([1, 0, 0, 0, 0, 0, 6, 1], np.uint64)
    assert np.array_equal(tot, np.stack([per_halo, per_halo]))
    _, _, etot, _ = t.parse_result(t.read_expected_text(case["expected"]), k)
    assert np.array_equal(etot, 3 * per_halo)


def test_dpps_equivalence_on_this_cpu():
    t.oracle()
    out = subprocess.run([t.DPPS_PROBE, "3000000", "12345"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("dpps ok")


def test_thresholds_match_reference_labels():
    """The %f radius labels of every golden result equal the oracle's
    thresholds printed the same way (outputResult: sqrt of r^2)."""
    for case in CASES:
        mn, mx, k = t.parse_radius(case["radius"])
        _, r2 = t.oracle_thresholds(mn, mx, k)
        *_, labels = t.parse_result(t.read_expected_text(case["expected"]), k)
        rad = [np.float32(np.sqrt(np.float64(v))) for v in r2]
        want = ["  [0.0 ~ %f]" % rad[0]] + ["  [%f ~ %f]" % (rad[i - 1], rad[i])
                                          for i in range(1, k)]
        assert labels == want, case["name"]


def test_fma_does_not_move_config_thresholds():
    """Logspace's a + i*c is contracted to an FMA by gcc -O2 on an FMA host;
    for every radius spec of the workloads both forms agree bit for bit, so
    the FMA is kept for exactness but cannot be observed in these counts."""
    for mn, mx, k in [(0.001, 5, 8), (0.001, 5, 40), (0.01, 20, 64), (0.5, 100, 16)]:
        r = np.zeros(k, np.float32)
        r2a = np.zeros(k, np.float32)
        r2b = np.zeros(k, np.float32)
        t.oracle().oracle_shell_thresholds(mn, mx, k, r.ctypes.data, r2a.ctypes.data)
        t.oracle().oracle_shell_thresholds_nofma(mn, mx, k, r.ctypes.data, r2b.ctypes.data)
        assert np.array_equal(r2a.view(np.uint32), r2b.view(np.uint32))


def lattice_counts(a_num, a_den, n, r2_units, center_offset):
    """Closed form: particles on the dyadic lattice c + a*(i, j, k), |i|,|j|,|k|
    <= n, a = a_num / a_den; all distances and box edges are exact in FP32, so
    the count in shell s is the number of integer triples with
    r2_units[s-1] <= i^2 + j^2 + k^2 < r2_units[s] and |i|,|j|,|k| <= rbox/a."""
    rbox_units = int(np.floor(np.sqrt(r2_units[-1])))
    counts = [0] * len(r2_units)
    for i, j, k in itertools.product(range(-n, n + 1), repeat=3):
        if max(abs(i), abs(j), abs(k)) > rbox_units:
            continue
        q = i * i + j * j + k * k
        for s, lim in enumerate(r2_units):
            if q < lim:
                counts[s] += 1
                break
    return counts


@pytest.mark.parametrize("r2_units", [[1, 2, 3, 4], [2, 5, 9, 16], [1, 9, 25], [3, 8, 12, 25]])
def test_lattice_closed_form(r2_units):
    a = 0.25
    n = 7
    c = np.array([300.0, 400.0, 500.0], np.float32)
    g = np.arange(-n, n + 1, dtype=np.float32) * np.float32(a)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + c
    r2 = np.array(r2_units, np.float32) * np.float32(a * a)
    got, _ = t.oracle_count(pts.astype(np.float32), c[None, :], r2)
    assert got[0].tolist() == lattice_counts(1, 4, n, r2_units, c)


def test_periodic_images_closed_form():
    """A halo at a box corner sees particles across all faces: with a
    lattice around (0,0,0) wrapped into [0,1000), every image is needed and
    the counts equal the un-wrapped lattice's closed form."""
    a = 0.5
    n = 4
    g = np.arange(-n, n + 1, dtype=np.float32) * np.float32(a)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    wrapped = np.where(pts < 0, pts + np.float32(1000.0), pts).astype(np.float32)
    r2_units = [1, 3, 6, 13]
    r2 = np.array(r2_units, np.float32) * np.float32(a * a)
    got, _ = t.oracle_count(wrapped, np.zeros((1, 3), np.float32), r2)
    assert got[0].tolist() == lattice_counts(1, 2, n, r2_units, None)


def test_images_order_and_shift_rule():
    """Image centres of a corner halo: all 8, +1000 on low-side axes, -1000
    on high-side axes; the low side wins when both hold (r_box > 500)."""
    o = t.oracle()
    cent = np.zeros((8, 3), np.float32)
    n = o.oracle_images(np.array([1.0, 999.0, 500.0], np.float32).ctypes.data,
                        np.float32(5.0), cent.ctypes.data)
    got = {tuple(map(float, r)) for r in cent[:n]}
    assert got == {(1.0, 999.0, 500.0), (1001.0, 999.0, 500.0), (1.0, -1.0, 500.0),
                   (1001.0, -1.0, 500.0)}
    n = o.oracle_images(np.array([500.0, 500.0, 500.0], np.float32).ctypes.data,
                        np.float32(600.0), cent.ctypes.data)
    assert n == 8 and cent[:n].max(axis=0).tolist() == [1500.0, 1500.0, 1500.0]
    assert cent[:n].min(axis=0).tolist() == [500.0, 500.0, 500.0]


# ------------------------------------------------- compressed (Gadget) decode --

from fractions import Fraction  # noqa: E402


def f32_nearest(x: Fraction) -> np.float32:
    """x rounded to the nearest FP32 value, ties to even, from exact x."""
    if x == 0:
        return np.float32(0.0)
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    while Fraction(2) ** e > a:
        e -= 1
    while Fraction(2) ** (e + 1) <= a:
        e += 1
    ulp = Fraction(2) ** max(e - 23, -149)
    q = a / ulp
    n = q.numerator // q.denominator
    rest = q - n
    if rest > Fraction(1, 2) or (rest == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return np.float32(float(n * ulp) if x > 0 else -float(n * ulp))


def lattice_point(pid, ng):
    """Closed form of a grid id 1..ng^3 with one tile: id - 1 counts the
    lattice with z fastest, then y, then x."""
    k = pid - 1
    return k // (ng * ng), (k // ng) % ng, k % ng


def exact_decode(pid, uv, ng, ci, dmin, unitsep, fused):
    """pos[d] = (float)((double)s + unitsep*uv[d]), s = ci*ip + dmin rounded
    once (fused) or twice, each rounding from the exact value."""
    out = []
    for d, ip in enumerate(lattice_point(pid, ng)):
        prod = Fraction(float(ci)) * ip
        if fused:
            s = f32_nearest(prod + Fraction(float(dmin)))
        else:
            s = f32_nearest(Fraction(float(f32_nearest(prod))) + Fraction(float(dmin)))
        v = float(Fraction(float(s)) + Fraction(unitsep) * int(uv[d]))
        out.append(np.float32(v))
    return out


def gadget_header(total, box, dmin, dmax, bits, id_start, npart=0):
    import struct
    h = bytearray(256)
    struct.pack_into("<6i", h, 0, 0, npart, 0, 0, 0, 0)
    lo = total & 0xFFFFFFFF
    lo = lo - (1 << 32) if lo >= (1 << 31) else lo
    struct.pack_into("<6i", h, 96, 0, lo, total >> 32, 0, 0, 0)
    struct.pack_into("<d", h, 128, box)
    struct.pack_into("<ffi", h, 172, dmin, dmax, bits)
    struct.pack_into("<q", h, 184, id_start)
    return bytes(h)


@pytest.mark.parametrize("ng,box,dmin,dmax,bits", [
    (24, 1000.0, -30.0, 30.0, 16), (48, 1000.0, -7.3, 7.3, 16), (1700, 1000.0, -2.0, 2.0, 16),
    (37, 687.5, -11.1, 3.3, 12), (10, 500.0, 5.0, -5.0, 10)])
def test_gadget_decode_exact_arithmetic(ng, box, dmin, dmax, bits):
    """The oracle's decode equals the decode computed from exact rationals
    with each of the reference's roundings applied once, fused and not, on
    ids spread over the whole grid; the two differ somewhere, so the check
    tells them apart."""
    rng = np.random.default_rng(ng)
    prm = t.gadget_params(gadget_header(ng ** 3, box, dmin, dmax, bits, 1))
    assert prm["ng1d"] == ng and prm["nt1d"] == 1
    assert prm["cell_interval"] == np.float32(box / ng)
    assert prm["unitsep"] == float(np.float32(np.float32(dmax) - np.float32(dmin))) / 2 ** bits
    ids = rng.integers(1, ng ** 3 + 1, 600)
    uv = rng.integers(0, 1 << bits, (600, 3), dtype=np.uint16)
    seen_diff = False
    for fused in (1, 0):
        got = np.stack([t.oracle_gadget_decode(uv[i], dict(prm, id_start=int(ids[i])), fused)[0]
                        for i in range(600)])
        want = np.array([exact_decode(int(ids[i]), uv[i], ng, prm["cell_interval"],
                                      prm["disp_min"], prm["unitsep"], fused)
                         for i in range(600)], np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), fused
        if fused:
            other = want
        else:
            seen_diff = not np.array_equal(other, want)
    if ng in (48, 1700, 37):  # grids where some ci*ip + dmin rounds differently
        assert seen_diff


def test_gadget_grid_closed_form():
    """Grid points of ids 1..ng^3 (one tile) are the lattice, z fastest; a
    second tile (ids past ng^3) continues along x; id 0 is (0, 0, -1)."""
    o = t.oracle()
    ip = np.zeros(3, np.int32)
    ng = 7
    pts = set()
    for pid in range(1, ng ** 3 + 1):
        o.oracle_gadget_gridpos(pid, ng, 1, ip.ctypes.data)
        assert tuple(ip) == lattice_point(pid, ng)
        pts.add(tuple(ip))
    assert len(pts) == ng ** 3
    o.oracle_gadget_gridpos(ng ** 3 + 1, ng, 1, ip.ctypes.data)
    assert tuple(ip) == (ng, 0, 0)
    o.oracle_gadget_gridpos(0, ng, 1, ip.ctypes.data)
    assert tuple(ip) == (0, 0, -1)
    # two tiles per side: tile t = (id-1) // ng^3 has digits (tx, ty, tz)
    # base 2, and shifts the lattice by ng per digit
    for pid in range(1, 8 * ng ** 3 + 1, 13):
        o.oracle_gadget_gridpos(pid, ng, 2, ip.ctypes.data)
        tile, inner = divmod(pid - 1, ng ** 3)
        g = lattice_point(inner + 1, ng)
        tdig = (tile // 4, (tile // 2) % 2, tile % 2)
        assert tuple(ip) == tuple(tdig[d] * ng + g[d] for d in range(3))


def test_gadget_params_totals():
    """nparticles_at1d is the cube root of npartTotal[1] + npartTotal[2]<<32
    (whole for exact cubes, also past 2^32), ng1d = at1d / ntile."""
    for b in (1, 2, 24, 1000, 1200, 1700, 2048, 3000):
        prm = t.gadget_params(gadget_header(b ** 3, 1000.0, 0.0, 1.0, 16, 1))
        assert prm["ng1d"] == b
        prm = t.gadget_params(gadget_header(b ** 3, 1000.0, 0.0, 1.0, 16, 1), ntile=2)
        assert prm["ng1d"] == b // 2 and prm["nt1d"] == 2
    prm = t.gadget_params(gadget_header(25 ** 3 - 1, 1000.0, 0.0, 1.0, 16, 1))
    assert prm["ng1d"] == 24  # rounds down
    # npartTotal[1] is an int there: totals in [2^31, 2^32) read as negative
    # and give no grid (the product refuses such a header)
    prm = t.gadget_params(gadget_header(1625 ** 3, 1000.0, 0.0, 1.0, 16, 1))
    assert prm["ng1d"] <= 0


def test_gadget_fma_case_tells_the_decodes_apart():
    """gad_fma's halos sit where the fused and the unfused decode put a
    particle in different shells: the reference's golden result equals the
    fused decode's counts (checked with every golden case) and differs from
    the unfused decode's in every row."""
    case = [c for c in CASES if c["name"] == "gad_fma"][0]
    mn, mx, k = t.parse_radius(case["radius"])
    _, r2 = t.oracle_thresholds(mn, mx, k)
    _, pts = t.read_halos(t.GOLD + "/" + case["halo_file"])
    _, ecnt, _, _ = t.parse_result(t.read_expected_text(case["expected"]), k)
    plain = t.oracle_decode_file(t.GOLD + "/" + case["particle_files"][0], fused=0)
    c0, _ = t.oracle_count(plain, pts, r2)
    assert pts.shape[0] >= 10
    assert (c0 != ecnt).any(axis=1).all()


# ------------------------------------------------ rt_rcount's R-tree count --

RT_CASES = [c for c in CASES if "rt_expected" in c]


def oracle_rt_case_counts(case):
    mn, mx, k = t.parse_radius(case["radius"])
    _, r2 = t.oracle_thresholds(mn, mx, k)
    chunks, ids, pts = t.case_inputs(case)
    tot = np.zeros((pts.shape[0], k), np.uint64)
    for ch in chunks:
        tot += t.oracle_rt_count(ch, pts, r2)
    return ids, tot, k


@pytest.mark.parametrize("case", RT_CASES, ids=[c["name"] for c in RT_CASES])
def test_rt_oracle_matches_reference_rt_rcount(case):
    """The R-tree restatement against rt_rcount's own results (oracle/_ref,
    built from the reference sources), cases where its counts differ from
    astr_rcount's among them."""
    ids, got, k = oracle_rt_case_counts(case)
    eids, ecnt, _, _ = t.parse_result(t.read_expected_text(case["rt_expected"]), k)
    assert np.array_equal(eids, ids)
    assert np.array_equal(got, ecnt)


def test_rt_fixtures_exercise_the_leaf_rule():
    """The rt fixtures hold rows where the two CLIs disagree (a halo on a
    particle), so they pin the leaf rule and not only the shared part."""
    differing = 0
    for case in RT_CASES:
        _, astr, k = oracle_case_counts(case)
        _, ecnt, _, _ = t.parse_result(t.read_expected_text(case["rt_expected"]), k)
        differing += int((astr != ecnt).any(axis=1).sum())
    assert differing >= 1500


def test_rt_equals_astr_off_the_particles():
    """With no halo on a particle (and no box-edge rounding case), the leaf
    rule never fires and an R-tree walk that prunes by exact boxes finds the
    pairs the brute force finds: periodic images included."""
    rng = np.random.default_rng(11)
    p = rng.random((30000, 3), dtype=np.float32) * np.float32(1000)
    h = rng.random((3000, 3), dtype=np.float32) * np.float32(1000)
    h[:300, 0] = rng.uniform(0, 8, 300).astype(np.float32)
    h[300:600, 1] = rng.uniform(992, 1000, 300).astype(np.float32)
    h[600:700, :] = rng.uniform(0, 8, (100, 3)).astype(np.float32)
    _, r2 = t.oracle_thresholds(0.5, 30.0, 12)
    want, _ = t.oracle_count(p, h, r2)
    assert want.sum() > 5000
    assert np.array_equal(t.oracle_rt_count(p, h, r2), want)


def test_rt_zero_distance_ends_the_leaf_scan():
    """Ten points fill the root leaf in insertion order; a halo on point 4
    counts points 0-3 only (src/objects/RTree/MultiRangeCount.c:104-105),
    where astr_rcount counts all ten, point 4 itself in the first shell."""
    xs = np.float32(500.0) + np.arange(10, dtype=np.float32) * np.float32(0.125)
    p = np.stack([xs, np.full(10, 500, np.float32), np.full(10, 500, np.float32)], 1)
    r2 = np.array([0.0625, 0.25, 1.0, 4.0], np.float32)  # r = 1/4, 1/2, 1, 2
    h = p[[4]]
    # distances 0.5, 0.375, 0.25, 0.125 -> squared .25, .140625, .0625, .015625
    want_rt = np.array([[1, 2, 1, 0]], np.uint64)
    with t.OracleRTree(p) as tr:
        assert tr.height() == 0
        got = tr.count(h, r2)
    assert np.array_equal(got, want_rt)
    astr, _ = t.oracle_count(p, h, r2)
    # astr: 0 (itself) and 0.125 -> shell 0; 0.25 (x2) -> shell 1 (not < .0625);
    # 0.375, 0.5 (x2 each side where present) -> shells 2, 2; 0.625 -> 2
    d2 = ((p[:, 0] - h[0, 0]) ** 2).astype(np.float32)
    want_astr = np.zeros((1, 4), np.uint64)
    for v in d2:
        want_astr[0, int(np.argmax(v < r2))] += 1 if (v < r2).any() else 0
    assert np.array_equal(astr, want_astr)
    assert astr.sum() == 10 and got.sum() == 4


def test_rt_tree_shape():
    """Structure the build must keep: every id once, leaf sizes within
    [MIN_ENT, MAX_ENT] below the root, the height of a balanced tree."""
    rng = np.random.default_rng(3)
    p = rng.random((50000, 3), dtype=np.float32) * np.float32(1000)
    with t.OracleRTree(p) as tr:
        ids, lens = tr.leaves()
        h = tr.height()
    assert np.array_equal(np.sort(ids), np.arange(p.shape[0], dtype=np.uint64))
    assert lens.sum() == p.shape[0]
    assert lens.min() >= 4 and lens.max() <= 10
    # at least 4 and at most 10 children per inner node
    assert np.ceil(np.log(lens.size) / np.log(10)) <= h <= np.floor(np.log(lens.size) / np.log(4)) + 1

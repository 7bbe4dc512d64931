"""Host side of the drop-in CLI, CPU only: argv errors (stderr, exit status,
stdout) against the reference's, the halo and filter readers, shell
thresholds, and the result writer byte for byte -- all without a GPU."""
import ctypes
import gzip
import os
import subprocess

import numpy as np
import pytest

import sjcs_testlib as t

MANIFEST = t.manifest()
CASES = MANIFEST["cases"]


class Halo(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("mass", ctypes.c_float),
                ("point", ctypes.c_float * 3)]


class Halos(ctypes.Structure):
    _fields_ = [("halo", ctypes.POINTER(Halo)), ("n", ctypes.c_uint64)]


@pytest.fixture(scope="module")
def host(built):
    lib = ctypes.CDLL(built.HOST_LIB_PATH)
    vp = ctypes.c_void_p
    lib.sjcs_read_halos.argtypes = [ctypes.c_char_p, ctypes.POINTER(Halos), ctypes.c_int,
                                    ctypes.c_char_p, ctypes.c_size_t]
    lib.sjcs_read_and_filter_halos.argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                               ctypes.POINTER(Halos), ctypes.c_char_p,
                                               ctypes.c_size_t]
    lib.sjcs_free_halos.argtypes = [ctypes.POINTER(Halos)]
    lib.sjcs_write_result.argtypes = [ctypes.c_char_p, ctypes.POINTER(Halos), vp, vp,
                                      ctypes.c_int, ctypes.c_int]
    lib.sjcs_host_shell_thresholds.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                               vp, vp]
    lib.sjcs_particle_count.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64)]
    lib.sjcs_parse_halo_line.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_float)]
    return lib


def read_with_host(host, path, nthreads=4, filter_path=None):
    hs = Halos()
    err = ctypes.create_string_buffer(4096)
    if filter_path is None:
        rc = host.sjcs_read_halos(path.encode(), ctypes.byref(hs), nthreads, err, 4096)
    else:
        rc = host.sjcs_read_and_filter_halos(path.encode(), filter_path.encode(),
                                             ctypes.byref(hs), err, 4096)
    if rc != 0:
        return None, err.value.decode()
    ids = np.array([hs.halo[i].id for i in range(hs.n)], np.uint64)
    pts = np.array([list(hs.halo[i].point) for i in range(hs.n)], np.float32).reshape(-1, 3)
    return (hs, ids, pts), ""


@pytest.mark.parametrize("i", range(len(MANIFEST["option_errors"])))
def test_option_errors_match_reference(built, i, tmp_path):
    case = MANIFEST["option_errors"][i]
    r = subprocess.run([built.CLI_PATH] + case["args"], capture_output=True, text=True,
                       cwd=tmp_path, timeout=60)
    assert r.returncode == case["returncode"]
    assert r.stderr.replace(built.CLI_PATH, "ARGV0") == case["stderr"]
    assert r.stdout == case["stdout"]


@pytest.mark.parametrize("nthreads", [1, 3, 16])
@pytest.mark.parametrize("name", ["uni_halos.list", "fmt_halos.list", "tie_halos.list",
                                  "tdpb_halos.list", "rtree_halos.list"])
def test_halo_reader_matches_libc_restatement(host, name, nthreads):
    path = os.path.join(t.GOLD, name)
    (hs, ids, pts), err = read_with_host(host, path, nthreads)
    want_ids, want_pts = t.read_halos(path)
    assert np.array_equal(ids, want_ids)
    assert np.array_equal(pts.view(np.uint32), want_pts.view(np.uint32))
    host.sjcs_free_halos(ctypes.byref(hs))


def test_halo_reader_edge_records(host, tmp_path):
    """fgets record rules: a line longer than 16383 bytes is several records,
    a last line without '\\n' is dropped, a NUL ends the parsed string."""
    long_tail = "9" * 20000
    lines = [f"# h {i}" for i in range(16)]
    body = ["1 0 5 0 0 0 0 0 1.5 2.5 3.5 0",
            "2 0 5 0 0 0 0 0 4 5 6 " + long_tail,
            "3 0 5 0 0 0 0 0 7\x008 9 10 11 12"]
    p = tmp_path / "h.list"
    p.write_bytes(("\n".join(lines + body) + "\n4 0 5 0 0 0 0 0 1 1 1").encode())
    res, err = read_with_host(host, str(p))
    # the tail of line 2 is its own record with no spaces: the reference fails
    assert res is None and "may be incorrect" in err
    try:
        t.read_halos(str(p))
        raise AssertionError("restatement accepted it")
    except ValueError:
        pass
    p.write_bytes(("\n".join(lines + body[:1]) + "\n4 0 5 0 0 0 0 0 1 1 1").encode())
    (hs, ids, pts), _ = read_with_host(host, str(p))
    assert ids.tolist() == [1] and pts.tolist() == [[1.5, 2.5, 3.5]]
    host.sjcs_free_halos(ctypes.byref(hs))


def test_halo_reader_errors(host, tmp_path):
    res, err = read_with_host(host, str(tmp_path / "missing.list"))
    assert res is None and err.startswith("ERROR: file ") and "not found in readHalos()" in err
    p = tmp_path / "short.list"
    p.write_text("# only\n# two headers\n")
    res, err = read_with_host(host, str(p))
    assert res is None and "does not contain headers in readHalos()" in err


def test_filter_reader_intended_semantics(host):
    path = os.path.join(t.GOLD, "uni_halos.list")
    fpath = os.path.join(t.GOLD, "uni_filter.txt")
    (hs, ids, pts), _ = read_with_host(host, path, filter_path=fpath)
    ids0, pts0 = t.read_halos(path)
    want_ids, want_pts = t.filter_halos(ids0, pts0, fpath)
    assert np.array_equal(ids, want_ids) and np.array_equal(pts, want_pts)
    # every listed id below 2000 is kept (ids are 0..1999, sorted, unique)
    listed = {int(x) for x in open(fpath).read().split()}
    assert set(ids.tolist()) == {i for i in listed if i < 2000}
    host.sjcs_free_halos(ctypes.byref(hs))


@pytest.mark.parametrize("spec", ["0.001:5:8", "0.001:5:40", "0.01:20:64", "5:0.001:8",
                                  "0.5:100:16", "3:7:1", "0.25:2:7", "1e-6:900:300"])
def test_thresholds_bitwise_equal_oracle(host, spec):
    mn, mx, k = t.parse_radius(spec)
    r = np.zeros(k, np.float32)
    r2 = np.zeros(k, np.float32)
    host.sjcs_host_shell_thresholds(mn, mx, k, r.ctypes.data, r2.ctypes.data)
    _, want = t.oracle_thresholds(mn, mx, k)
    assert np.array_equal(r2.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_result_writer_byte_identical(host, case, tmp_path):
    """Counts from the oracle (pinned to the reference) written by the host
    writer give the reference's result file byte for byte."""
    if case.get("filter_file"):
        pytest.skip("the reference's filter reads an uninitialised count; see test_cli")
    mn, mx, k = t.parse_radius(case["radius"])
    _, r2 = t.oracle_thresholds(mn, mx, k)
    chunks, ids, pts = t.case_inputs(case)
    tot = np.zeros((pts.shape[0], k), np.uint64)
    for ch in chunks:
        tot += t.oracle_count(ch, pts, r2)[0]
    arr = (Halo * max(1, len(ids)))()
    for i in range(len(ids)):
        arr[i].id = int(ids[i])
        arr[i].point[:] = pts[i].tolist()
    hs = Halos(ctypes.cast(arr, ctypes.POINTER(Halo)), len(ids))
    out = tmp_path / "out.txt"
    for nthreads in (1, 5):
        rc = host.sjcs_write_result(str(out).encode(), ctypes.byref(hs), tot.ctypes.data,
                                    r2.ctypes.data, k, nthreads)
        assert rc == 0
        assert out.read_bytes() == t.read_expected_text(case["expected"])


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_plain_formatter_equals_reference_files(case):
    """The plain formatter the large writer tests compare with gives every
    golden result file (the reference's own output) byte for byte from the
    oracle's counts: labels, %lu and the Total block."""
    if case.get("filter_file"):
        pytest.skip("the reference's filter reads an uninitialised count; see test_cli")
    mn, mx, k = t.parse_radius(case["radius"])
    _, r2 = t.oracle_thresholds(mn, mx, k)
    chunks, ids, pts = t.case_inputs(case)
    tot = np.zeros((pts.shape[0], k), np.uint64)
    for ch in chunks:
        tot += t.oracle_count(ch, pts, r2)[0]
    assert t.format_result(ids, tot, r2) == t.read_expected_text(case["expected"])


@pytest.mark.parametrize("spec,nthreads,top", [("0.001:5:40", 1, 1 << 20),
                                               ("0.001:5:40", 16, 1 << 20),
                                               ("0.01:20:64", 16, 1 << 20), ("2:3:1", 7, 1 << 20),
                                               ("0.001:5:40", 16, 150),
                                               ("100000:900000:3", 4, 150)])
def test_result_writer_many_batches(host, tmp_path, spec, nthreads, top):
    """140000 halos -- five of the writer's 32768-row batches, the last one
    partial -- with ids and counts over the whole u64 range (sums that wrap
    in the Total block included), or counts mostly of one and two digits
    (`top` 150), with labels of at most 32 bytes (the slot table) or longer
    (radii of six digits), written by the host writer with 1 to 16 threads:
    byte for byte the plain formatter's text."""
    rng = np.random.default_rng(len(spec) * 31 + nthreads + top)
    mn, mx, k = t.parse_radius(spec)
    _, r2 = t.oracle_thresholds(mn, mx, k)
    n = 140_000
    ids = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    ids[:3] = [0, 1, np.iinfo(np.uint64).max]
    counts = rng.integers(0, top, (n, k), dtype=np.uint64)
    counts[rng.random((n, k)) < 0.001] = np.iinfo(np.uint64).max - np.uint64(5)
    counts[0, 0] = np.iinfo(np.uint64).max  # halo 0 is doubled in the Total block
    arr = (Halo * n)()
    ctypes.memmove(arr, np.rec.fromarrays([ids, np.zeros(n, np.float32)] +
                                          [np.zeros(n, np.float32)] * 3,
                                          dtype=[("id", "<u8"), ("m", "<f4"), ("x", "<f4"),
                                                 ("y", "<f4"), ("z", "<f4")]).tobytes(),
                   ctypes.sizeof(Halo) * n)
    hs = Halos(ctypes.cast(arr, ctypes.POINTER(Halo)), n)
    out = tmp_path / "out.txt"
    rc = host.sjcs_write_result(str(out).encode(), ctypes.byref(hs), counts.ctypes.data,
                                r2.ctypes.data, k, nthreads)
    assert rc == 0
    got = out.read_bytes()
    want = t.format_result(ids, counts, r2)
    assert got == want


ROWS32_FN = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64)


@pytest.mark.parametrize("spec,nthreads,top", [("0.001:5:40", 16, 150),
                                               ("0.01:20:64", 3, 1 << 32),
                                               ("100000:900000:3", 5, 1 << 32)])
def test_result_writer_rows32(host, tmp_path, spec, nthreads, top):
    """The writer fed u32 count rows through its callback (batch by batch,
    the CLI's read-back of 32-bit totals) writes the same bytes as the u64
    writer on the same counts: 140000 halos, small counts or counts up to
    2^32 - 1, slot-table labels and longer ones."""
    rng = np.random.default_rng(len(spec) + nthreads)
    mn, mx, k = t.parse_radius(spec)
    _, r2 = t.oracle_thresholds(mn, mx, k)
    n = 140_000
    ids = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    c32 = rng.integers(0, top, (n, k), dtype=np.uint64).astype(np.uint32)
    c32[0, 0] = np.iinfo(np.uint32).max
    c64 = c32.astype(np.uint64)
    arr = (Halo * n)()
    ctypes.memmove(arr, np.rec.fromarrays([ids, np.zeros(n, np.float32)] +
                                          [np.zeros(n, np.float32)] * 3,
                                          dtype=[("id", "<u8"), ("m", "<f4"), ("x", "<f4"),
                                                 ("y", "<f4"), ("z", "<f4")]).tobytes(),
                   ctypes.sizeof(Halo) * n)
    hs = Halos(ctypes.cast(arr, ctypes.POINTER(Halo)), n)
    asked = []

    def rows(ctx, beg, cnt):
        asked.append((beg, cnt))
        return c32.ctypes.data + int(beg) * k * 4

    cb = ROWS32_FN(rows)
    host.sjcs_write_result_rows32.argtypes = [ctypes.c_char_p, ctypes.POINTER(Halos), ROWS32_FN,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                              ctypes.c_int]
    out32, out64 = tmp_path / "r32.txt", tmp_path / "r64.txt"
    assert host.sjcs_write_result_rows32(str(out32).encode(), ctypes.byref(hs), cb, None,
                                         r2.ctypes.data, k, nthreads) == 0
    assert host.sjcs_write_result(str(out64).encode(), ctypes.byref(hs), c64.ctypes.data,
                                  r2.ctypes.data, k, nthreads) == 0
    assert out32.read_bytes() == out64.read_bytes()
    assert [a for a, _ in asked] == list(range(0, n, 32768))


def test_particle_count_ignores_trailing_bytes(host, tmp_path):
    p = tmp_path / "p.dat"
    p.write_bytes(np.arange(30, dtype=np.float32).tobytes() + b"\x01\x02\x03\x04\x05")
    n = ctypes.c_uint64()
    assert host.sjcs_particle_count(str(p).encode(), ctypes.byref(n)) == 0
    assert n.value == 10
    assert host.sjcs_particle_count(str(tmp_path / "none").encode(), ctypes.byref(n)) == -1


def test_cli_missing_files_fail_cleanly(built, tmp_path):
    r = subprocess.run([built.CLI_PATH, "--thread", "2", "--next-setup", "1",
                        "--particle-files", "nope.dat", "--halo-file", "nope.list",
                        "--radius", "0.1:5:4"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=60)
    assert r.returncode == 1
    assert "ERROR: file nope.list not found in readHalos()." in r.stderr


# ---------------------------------------------- compressed (Gadget) files --

GADGET_FILES = sorted({p for c in CASES if "--compressed" in c["extra"]
                       for p in c["particle_files"]})


def host_read_compressed(built, path):
    lib = ctypes.CDLL(built.HOST_LIB_PATH)
    vp = ctypes.c_void_p
    lib.sjcs_read_compressed.argtypes = [ctypes.c_char_p, vp, ctypes.POINTER(vp),
                                         ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_uint64)]
    hdr = ctypes.create_string_buffer(256)
    uv = vp()
    cap, n = ctypes.c_uint64(0), ctypes.c_uint64(0)
    rc = lib.sjcs_read_compressed(path.encode(), hdr, ctypes.byref(uv), ctypes.byref(cap),
                                  ctypes.byref(n))
    out = None
    if rc == 0 and n.value:
        out = np.ctypeslib.as_array(ctypes.cast(uv, ctypes.POINTER(ctypes.c_uint16)),
                                    (n.value * 3,)).reshape(-1, 3).copy()
    libc = ctypes.CDLL("libc.so.6")
    libc.free.argtypes = [vp]
    libc.free(uv)
    return rc, hdr.raw, out


def test_compressed_reader(built):
    assert len(GADGET_FILES) >= 6
    for name in GADGET_FILES:
        path = os.path.join(t.GOLD, name)
        rc, hdr, uv = host_read_compressed(built, path)
        want_hdr, want_uv = t.read_compressed(path)
        assert rc == 0 and hdr == want_hdr
        assert np.array_equal(uv, want_uv)
        h2, uv2 = built.read_compressed(path)
        assert h2 == want_hdr and np.array_equal(uv2, want_uv)


def test_compressed_reader_errors(built, tmp_path):
    good = open(os.path.join(t.GOLD, "gad_grid.gdt"), "rb").read()
    assert host_read_compressed(built, str(tmp_path / "none"))[0] == -1
    short = tmp_path / "short.gdt"
    short.write_bytes(good[:-100])
    assert host_read_compressed(built, str(short))[0] == -2
    neg = bytearray(good)
    neg[8:12] = (-5).to_bytes(4, "little", signed=True)
    bad = tmp_path / "neg.gdt"
    bad.write_bytes(bytes(neg))
    assert host_read_compressed(built, str(bad))[0] == -2
    empty = bytearray(good[:268])
    empty[8:12] = (0).to_bytes(4, "little")
    e = tmp_path / "empty.gdt"
    e.write_bytes(bytes(empty))
    rc, _, uv = host_read_compressed(built, str(e))
    assert rc == 0 and uv is None


def test_gadget_decode_params_equal_oracle(built):
    """The library's header -> decode parameters (host code, no device)
    equal the oracle's on every golden file and on ntile 2."""
    from test_oracle import gadget_header
    headers = [t.read_compressed(os.path.join(t.GOLD, n))[0] for n in GADGET_FILES]
    headers += [gadget_header(2048 ** 3, 1000.0, -1.5, 2.5, 14, 3 * 10 ** 9),
                gadget_header(1700 ** 3, 750.0, 3.0, -3.0, 8, -7)]
    for hdr in headers:
        for ntile in (1, 2):
            got = built.gadget_decode_params(hdr, ntile)
            want = t.gadget_params(hdr, ntile)
            assert got.fused == 1
            for f in ("id_start", "ng1d", "nt1d", "cell_interval", "disp_min", "unitsep"):
                assert getattr(got, f) == want[f], f
    with pytest.raises(built.SJCSError):
        built.gadget_decode_params(gadget_header(1625 ** 3, 1000.0, 0.0, 1.0, 16, 1))


def float_forms(rng):
    """Decimal spellings a halo file may hold, and those that sit exactly on
    a float midpoint or outside the fast path."""
    out = []
    f32 = rng.random(3000).astype(np.float32) * np.float32(1000.0)
    wide = (rng.standard_normal(3000) * 10.0 ** rng.integers(-30, 30, 3000)).astype(np.float32)
    for v in np.concatenate([f32, wide]):
        v = float(v)
        for fmt in ("%f", "%.9g", "%.17g", "%e", "%.3f", "%.12f", "%.20e"):
            out.append(fmt % v)
    # float midpoints written exactly: integers (ulp >= 2) and .5/.25 steps
    for _ in range(2000):
        e = int(rng.integers(21, 45))
        m = int(rng.integers(1 << 23, 1 << 24))
        ulp = 2 ** (e - 23)
        mid = m * ulp + ulp / 2
        out.append("%.3f" % mid if ulp < 2 else "%d" % int(mid))
    # long mantissas, large exponents, special forms
    for _ in range(2000):
        digits = "".join(rng.choice(list("0123456789"), int(rng.integers(1, 26))))
        dot = int(rng.integers(0, len(digits) + 1))
        s = digits[:dot] + "." + digits[dot:]
        if rng.random() < 0.5:
            s += "e%d" % int(rng.integers(-45, 45))
        out.append(("-" if rng.random() < 0.3 else "") + s)
    # significands around 2^53 (exact doubles stop there)
    for _ in range(3000):
        m = int(rng.integers(2 ** 52, 2 ** 60))
        e = int(rng.integers(-22, 23))
        out.append("%de%d" % (m, e))
    # significands above 2^53 next to float midpoints: rounding the
    # significand to double first would land on the wrong side
    out += ["606177124023437500e-14", "606177124023437501e-14", "606177124023437503e-14",
            "6033258007812499997e-14", "6033258007812499999e-14", "6033258007812500000e-14"]
    out += ["-0.0", "+0", "-0e5", "0x1.2cp+8", "-0X1p-3", "inf", "-nan", " 5.5", ".5", "5.",
            "1e-40", "1e-45", "3.4028235e38", "3.4028236e38", "3.5e38", "1e39", "1.5e", "2.5e+",
            "7.25abc", "1.17549435e-38", "1.1754942e-38", "9999999999999999999.5",
            "00000000000000000000000123.5", "16777217", "33554435", "0.000000000000000000001"]
    return out


def test_halo_line_numbers_equal_libc(host, tmp_path):
    """Every coordinate and mass the halo parser returns equals the C
    library's strtof of the same text, bit for bit, across plain decimals,
    exact float midpoints, long mantissas, large exponents and the forms
    that only the library handles; ids equal strtoul."""
    rng = np.random.default_rng(5)
    forms = float_forms(rng)
    idv = ctypes.c_uint64()
    mass = ctypes.c_float()
    pt = (ctypes.c_float * 3)()
    bad = []
    for i, s in enumerate(forms):
        a, b, c = s, forms[(i * 7 + 3) % len(forms)], forms[(i * 13 + 5) % len(forms)]
        ids = ["%d" % (i * 7919), "18446744073709551615", "99999999999999999999", "-3", " 42"][i % 5]
        line = f"{ids} 0 {b} 0 0 0 0 0 {a} {b} {c} 0\n".encode()
        ref = t.parse_halo_line(line)
        rc = host.sjcs_parse_halo_line(line, ctypes.byref(idv), ctypes.byref(mass), pt)
        if ref is None:
            assert rc == -1
            continue
        assert rc == 0
        want = [ref[1]] + list(ref[2])
        got = [mass.value, pt[0], pt[1], pt[2]]
        wb = np.array(want, np.float32).view(np.uint32)
        gb = np.array(got, np.float32).view(np.uint32)
        if not np.array_equal(wb, gb) or idv.value != ref[0]:
            bad.append((a, b, c, ids))
    assert bad == [], bad[:5]

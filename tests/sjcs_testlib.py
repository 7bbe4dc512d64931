"""Helpers shared by the tests: the oracle's ctypes bindings, readers of the
golden fixtures, and an independent restatement of the reference's halo
line parse (src/Gadget/BufHalo.c:218-249) that calls the C library's own
strtoul/strtof through ctypes, so the oracle tests do not depend on the
product's parser."""
from __future__ import annotations

import ctypes
import gzip
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_lib", "libsjcs_oracle.so")
DPPS_PROBE = os.path.join(ORACLE_DIR, "_lib", "dpps_probe")

_libc = ctypes.CDLL("libc.so.6")
_libc.strtoul.restype = ctypes.c_ulong
_libc.strtoul.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int]
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
_libc.strtol.restype = ctypes.c_long
_libc.strtol.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int]


def manifest():
    with open(os.path.join(GOLD, "manifest.json")) as f:
        return json.load(f)


# ------------------------------------------------------------------ oracle --

_oracle = None


def oracle():
    """The oracle library, built with gcc on first use if needed."""
    global _oracle
    if _oracle is not None:
        return _oracle
    if not os.path.exists(ORACLE_LIB) or not os.path.exists(DPPS_PROBE):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    lib = ctypes.CDLL(ORACLE_LIB)
    vp = ctypes.c_void_p
    lib.oracle_count.restype = ctypes.c_uint64
    lib.oracle_count.argtypes = [vp, ctypes.c_uint64, vp, ctypes.c_size_t, ctypes.c_uint64,
                                 vp, ctypes.c_int, vp, ctypes.c_int]
    lib.oracle_shell_thresholds.restype = None
    lib.oracle_shell_thresholds.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                            vp, vp]
    lib.oracle_shell_thresholds_nofma.restype = None
    lib.oracle_shell_thresholds_nofma.argtypes = lib.oracle_shell_thresholds.argtypes
    lib.oracle_distance.restype = ctypes.c_float
    lib.oracle_distance.argtypes = [vp, vp]
    lib.oracle_shell_of.restype = ctypes.c_int
    lib.oracle_shell_of.argtypes = [ctypes.c_float, vp, ctypes.c_int]
    lib.oracle_images.restype = ctypes.c_int
    lib.oracle_images.argtypes = [vp, ctypes.c_float, vp]
    lib.oracle_box_radius.restype = ctypes.c_float
    lib.oracle_box_radius.argtypes = [vp, ctypes.c_int]
    ll = ctypes.c_longlong
    lib.oracle_gadget_params.restype = None
    lib.oracle_gadget_params.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp, vp]
    lib.oracle_gadget_gridpos.restype = None
    lib.oracle_gadget_gridpos.argtypes = [ll, ll, ctypes.c_int, vp]
    lib.oracle_gadget_decode.restype = None
    lib.oracle_gadget_decode.argtypes = [vp, ctypes.c_uint64, ll, ll, ctypes.c_int,
                                         ctypes.c_float, ctypes.c_float, ctypes.c_double,
                                         ctypes.c_int, vp]
    lib.oracle_rt_build.restype = vp
    lib.oracle_rt_build.argtypes = [vp, ctypes.c_uint64]
    lib.oracle_rt_free.restype = None
    lib.oracle_rt_free.argtypes = [vp]
    lib.oracle_rt_height.restype = ctypes.c_int
    lib.oracle_rt_height.argtypes = [vp]
    lib.oracle_rt_leaves.restype = ctypes.c_uint64
    lib.oracle_rt_leaves.argtypes = [vp, vp, vp]
    lib.oracle_rt_count.restype = None
    lib.oracle_rt_count.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_uint64, vp, ctypes.c_int,
                                    vp]
    _oracle = lib
    return lib


def oracle_thresholds(minrad, maxrad, k):
    r = np.zeros(k, np.float32)
    r2 = np.zeros(k, np.float32)
    oracle().oracle_shell_thresholds(minrad, maxrad, k, r.ctypes.data, r2.ctypes.data)
    return r, r2


def oracle_count(particles, halos, r2, threads=8):
    """(counts [nh][K] uint64, W) by brute force."""
    p = np.ascontiguousarray(particles, np.float32).reshape(-1, 3)
    h = np.ascontiguousarray(halos, np.float32).reshape(-1, 3)
    r2 = np.ascontiguousarray(r2, np.float32)
    c = np.zeros((h.shape[0], r2.size), np.uint64)
    w = oracle().oracle_count(p.ctypes.data, p.shape[0], h.ctypes.data, 12, h.shape[0],
                              r2.ctypes.data, r2.size, c.ctypes.data, threads)
    return c, w


class OracleRTree:
    """rt_rcount's R-tree over `particles`, inserted in order (the oracle's
    restatement of src/objects/RTree/Build.c)."""

    def __init__(self, particles):
        self.p = np.ascontiguousarray(particles, np.float32).reshape(-1, 3)
        self.t = oracle().oracle_rt_build(self.p.ctypes.data, self.p.shape[0])
        if not self.t:
            raise ValueError("oracle R-tree build failed")

    def close(self):
        if self.t:
            oracle().oracle_rt_free(self.t)
            self.t = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def height(self):
        return oracle().oracle_rt_height(self.t)

    def leaves(self):
        """(ids in leaf order, leaf lengths), leaves depth first."""
        n = self.p.shape[0]
        ids = np.zeros(max(n, 1), np.uint64)
        lens = np.zeros(max(n, 1), np.uint32)
        nl = oracle().oracle_rt_leaves(self.t, ids.ctypes.data, lens.ctypes.data)
        return ids[:n], lens[:nl]

    def count(self, halos, r2):
        h = np.ascontiguousarray(halos, np.float32).reshape(-1, 3)
        r2 = np.ascontiguousarray(r2, np.float32)
        c = np.zeros((h.shape[0], r2.size), np.uint64)
        oracle().oracle_rt_count(self.t, h.ctypes.data, 12, h.shape[0], r2.ctypes.data,
                                 r2.size, c.ctypes.data)
        return c


def oracle_rt_count(particles, halos, r2):
    """rt_rcount's counts [nh][K] for one particle file."""
    with OracleRTree(particles) as t:
        return t.count(halos, r2)


def shell_index(dist, r2):
    """First shell k with dist < r2[k], or -1 (MultiRangeCount.c:107-113)."""
    r2 = np.ascontiguousarray(r2, np.float32)
    return oracle().oracle_shell_of(float(dist), r2.ctypes.data, r2.size)


def gadget_params(hdr, ntile=1):
    """The oracle's decode parameters of a 256-byte header, as a dict."""
    lib = oracle()
    h = np.frombuffer(bytes(hdr), np.uint8).copy()
    ids, ng = ctypes.c_longlong(), ctypes.c_longlong()
    nt = ctypes.c_int()
    ci, dmin, us = ctypes.c_float(), ctypes.c_float(), ctypes.c_double()
    lib.oracle_gadget_params(h.ctypes.data, ntile, ctypes.byref(ids), ctypes.byref(ng),
                             ctypes.byref(nt), ctypes.byref(ci), ctypes.byref(dmin),
                             ctypes.byref(us))
    return {"id_start": ids.value, "ng1d": ng.value, "nt1d": nt.value,
            "cell_interval": ci.value, "disp_min": dmin.value, "unitsep": us.value}


def oracle_gadget_decode(uv, prm, fused=1):
    uv = np.ascontiguousarray(uv, np.uint16).reshape(-1, 3)
    out = np.zeros(uv.shape, np.float32)
    oracle().oracle_gadget_decode(uv.ctypes.data, uv.shape[0], prm["id_start"], prm["ng1d"],
                                  prm["nt1d"], prm["cell_interval"], prm["disp_min"],
                                  prm["unitsep"], fused, out.ctypes.data)
    return out


def read_compressed(path):
    """(256 header bytes, uint16 triplets [n][3]) of a compressed particle
    file: skip 4, header, skip 8, npart[1] triplets (BufVector3D.c:111-136)."""
    with open(path, "rb") as f:
        raw = f.read()
    hdr = raw[4:260]
    n = int(np.frombuffer(hdr[4:8], "<i4")[0])
    uv = np.frombuffer(raw[268:268 + 6 * n], "<u2").reshape(n, 3)
    return hdr, uv


def oracle_decode_file(path, fused=1):
    hdr, uv = read_compressed(path)
    return oracle_gadget_decode(uv, gadget_params(hdr), fused)


def parse_radius(spec):
    """`--radius` as astr_rcount reads it (src/astr_rcount/main.c:506-533)."""
    mn = float(spec.split(":")[0])
    rest = spec.split(":", 1)[1]
    mx = float(_strtod_prefix(rest))
    k = int(_libc.strtol(spec.rsplit(":", 1)[1].encode(), None, 10))
    if mn > mx:
        mn, mx = mx, mn
    return mn, mx, k


def _strtod_prefix(s):
    strtod = _libc.strtod
    strtod.restype = ctypes.c_double
    strtod.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    return strtod(s.encode(), None)


# ----------------------------------------------------------------- inputs --

def read_particles(path):
    raw = np.fromfile(path, dtype=np.uint8)
    n = raw.size // 12
    return raw[: n * 12].view(np.float32).reshape(n, 3)


def _fgets_records(data: bytes, cap=16383):
    """Yields (record, at_eof) as fgets(buf, 16384) + feof would."""
    pos = 0
    size = len(data)
    while True:
        left = size - pos
        lim = min(left, cap)
        nl = data.find(b"\n", pos, pos + lim)
        if nl >= 0:
            yield data[pos:nl + 1], False
            pos = nl + 1
        else:
            yield data[pos:pos + lim], left < cap
            if left < cap:
                return
            pos += lim


def parse_halo_line(rec: bytes):
    """GetDataFromLineFast with libc strtoul/strtof; None when a space is
    missing.  The record is cut at its first NUL, as a C string is."""
    s = rec.split(b"\0", 1)[0]
    hid = _libc.strtoul(s, None, 10)
    pos = 0
    for _ in range(2):
        i = s.find(b" ", pos)
        if i < 0:
            return None
        pos = i + 1
    mass = _libc.strtof(s[pos:], None)
    for _ in range(6):
        i = s.find(b" ", pos)
        if i < 0:
            return None
        pos = i + 1
    xyz = [_libc.strtof(s[pos:], None)]
    for _ in range(2):
        i = s.find(b" ", pos)
        if i < 0:
            return None
        pos = i + 1
        xyz.append(_libc.strtof(s[pos:], None))
    return hid, mass, xyz


def read_halos(path):
    """(ids uint64, points float32 [n][3]) as readHalos reads the file."""
    with open(path, "rb") as f:
        data = f.read()
    recs = _fgets_records(data)
    for _ in range(16):
        rec, eof = next(recs)
        if eof:
            raise ValueError("no headers")
    ids, pts = [], []
    for rec, eof in recs:
        if eof:
            break
        r = parse_halo_line(rec)
        if r is None:
            raise ValueError("bad line")
        ids.append(r[0])
        pts.append(r[2])
    return np.array(ids, np.uint64), np.array(pts, np.float32).reshape(-1, 3)


def filter_halos(ids, pts, filter_path):
    """readAndFilterHalos' merge join (src/Gadget/BufHalo.c:121-214)."""
    with open(filter_path, "rb") as f:
        data = f.read()
    fl = [np.uint64(_libc.strtol(rec.split(b"\0", 1)[0], None, 10) & 0xFFFFFFFFFFFFFFFF)
          for rec, eof in _fgets_records(data) if rec]
    fl = sorted(int(v) for v in fl)
    keep = []
    fr = 0
    for i, hid in enumerate(ids.tolist()):
        if hid < fl[fr]:
            continue
        while hid > fl[fr]:
            fr += 1
            if fr >= len(fl):
                break
        if fr >= len(fl):
            break
        if hid == fl[fr]:
            keep.append(i)
    return ids[keep], pts[keep]


def filter_halos_nfilters_zero(ids, pts, filter_path):
    """What the reference actually does: ConstructFilterIdList
    (src/Gadget/BufHalo.c:251-300) never stores the id count, so
    fil->nfilters is whatever the fresh heap block holds -- 0 in practice.
    The list is then not sorted (RadixSort of 0 ids) and the merge join
    ends at the first halo whose id exceeds the file's first id."""
    with open(filter_path, "rb") as f:
        first = _libc.strtol(f.readline(), None, 10) & 0xFFFFFFFFFFFFFFFF
    keep = []
    for i, hid in enumerate(ids.tolist()):
        if hid < first:
            continue
        if hid > first:
            break
        keep.append(i)
    return ids[keep], pts[keep]


def case_inputs(case, filter_as_reference_runs=False):
    """Particle chunks, halo ids and points of a golden case."""
    if "--compressed" in case.get("extra", []):
        chunks = [oracle_decode_file(os.path.join(GOLD, p)) for p in case["particle_files"]]
    else:
        chunks = [read_particles(os.path.join(GOLD, p)) for p in case["particle_files"]]
    ids, pts = read_halos(os.path.join(GOLD, case["halo_file"]))
    if case.get("filter_file"):
        fn = filter_halos_nfilters_zero if filter_as_reference_runs else filter_halos
        ids, pts = fn(ids, pts, os.path.join(GOLD, case["filter_file"]))
    return chunks, ids, pts


# ---------------------------------------------------------------- results --

def read_expected_text(name):
    with gzip.open(os.path.join(GOLD, name), "rb") as f:
        return f.read()


def parse_result(text: bytes, k: int):
    """(ids, counts [n][K], total [K], labels [K]) from a result file."""
    lines = text.decode().split("\n")
    ids, rows = [], []
    i = 0
    labels = None
    while i < len(lines) and lines[i].startswith("id "):
        ids.append(int(lines[i][3:]))
        row, lab = [], []
        for j in range(k):
            head, val = lines[i + 1 + j].rsplit(": ", 1)
            row.append(int(val))
            lab.append(head)
        labels = lab
        rows.append(row)
        i += k + 1
    assert lines[i] == "" and lines[i + 1] == "Total", lines[i:i + 2]
    total, tlab = [], []
    for j in range(k):
        head, val = lines[i + 2 + j].rsplit(": ", 1)
        total.append(int(val))
        tlab.append(head)
    return (np.array(ids, np.uint64), np.array(rows, np.uint64).reshape(-1, k),
            np.array(total, np.uint64), tlab)


def format_result(ids, counts, r2):
    """The result text as outputResult writes it (src/astr_rcount/main.c:
    371-442), restated in plain Python for the writer tests: labels from
    the FP32 sqrt of each squared threshold printed with %f, "id %lu" and
    one line per shell, then the Total block, whose running sums live in
    halo 0's row (total = &counts[0], main.c:400): halo 0's counts are
    printed first and then added to themselves, so Total = 2 x halo 0 +
    the other halos, in u64 arithmetic."""
    k = len(r2)
    rad = [float(np.float32(np.sqrt(np.float64(v)))) for v in r2]
    labels = ["  [0.0 ~ %f]: " % rad[0]] + ["  [%f ~ %f]: " % (rad[i - 1], rad[i])
                                          for i in range(1, k)]
    counts = np.asarray(counts, np.uint64).reshape(-1, k)
    out = []
    for hid, row in zip(np.asarray(ids, np.uint64).tolist(), counts.tolist()):
        out.append("id %d\n" % hid)
        out.extend(lab + "%d\n" % v for lab, v in zip(labels, row))
    if counts.shape[0]:
        with np.errstate(over="ignore"):
            total = counts[0] + counts.sum(axis=0, dtype=np.uint64)
    else:
        total = np.zeros(k, np.uint64)
    out.append("\nTotal\n")
    out.extend(lab + "%d\n" % v for lab, v in zip(labels, total.tolist()))
    return "".join(out).encode()

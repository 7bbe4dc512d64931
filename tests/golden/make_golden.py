#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ by running the reference's
own astr_rcount (compiled from its sources by `make -f oracle/ref.mk` into
oracle/_ref/) on small inputs generated here.

Run from the repository root, in the container that has the reference:

    make -f oracle/ref.mk && python3 tests/golden/make_golden.py

Every fixture is data: input files, the reference's result file (gzipped)
and, for option errors, its stderr and exit status.  manifest.json lists
the cases.  The tests never run the reference; they compare against these
files.
"""
from __future__ import annotations

import gzip
import json
import os
import re
import shutil
import subprocess
import struct
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
GOLD = os.path.join(ROOT, "tests", "golden")
REF = os.path.join(ROOT, "oracle", "_ref", "astr_rcount")
RT_REF = os.path.join(ROOT, "oracle", "_ref", "rt_rcount")
RTREE_TEST = "/root/reference/include/rtree_test/particles.dat"


def write_halos(path, pts, ids=None, fmt="%f", trailing_newline=True, raw_lines=None):
    with open(path, "w") as f:
        for i in range(16):
            f.write(f"# dummy header {i}\n")
        lines = []
        for i, p in enumerate(pts):
            hid = i if ids is None else ids[i]
            c = " ".join(fmt % float(v) for v in p)
            lines.append(f"{hid} 0 1e+12 0 0 0 0 0 {c} 0 0 0 0 0")
        if raw_lines:
            lines.extend(raw_lines)
        f.write("\n".join(lines))
        if trailing_newline:
            f.write("\n")


def write_particles(path, pts):
    np.ascontiguousarray(pts, dtype=np.float32).tofile(path)


def write_compressed(path, uv, npart_total, box=1000.0, disp_min=-30.0, disp_max=30.0,
                     pos_bits=16, id_start=1):
    """A compressed (Gadget) particle file as readCompressedParticles reads it
    (src/Gadget/BufVector3D.c:111-136): record marker, 256-byte header
    (include/Gadget/GadgetHeader.h:4-27), two markers, uint16 triplets,
    marker."""
    uv = np.ascontiguousarray(uv, dtype="<u2").reshape(-1, 3)
    n = uv.shape[0]
    hdr = bytearray(256)
    struct.pack_into("<6i", hdr, 0, 0, n, 0, 0, 0, 0)                    # npart
    lo = npart_total & 0xFFFFFFFF
    lo = lo - (1 << 32) if lo >= (1 << 31) else lo
    struct.pack_into("<6i", hdr, 96, 0, lo, npart_total >> 32, 0, 0, 0)  # npartTotal
    struct.pack_into("<i", hdr, 124, 1)                                  # num_files
    struct.pack_into("<d", hdr, 128, box)                                # BoxSize
    struct.pack_into("<ffi", hdr, 172, disp_min, disp_max, pos_bits)
    struct.pack_into("<qq", hdr, 184, id_start, id_start + n - 1)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", 256) + bytes(hdr) + struct.pack("<i", 256))
        f.write(struct.pack("<i", 6 * n) + uv.tobytes() + struct.pack("<i", 6 * n))


def fma_sensitive_halos(path_particles, r2, shell, limit=40):
    """Halo centres that tell the two decodes apart: for particles whose
    coordinate on exactly one axis differs between cell_interval*ip +
    disp_min fused (one FMA) and unfused (two roundings), a centre on the
    particle's other two coordinates whose squared distance along that axis
    falls in `shell` for one decode and not for the other."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import sjcs_testlib as tl
    fused = tl.oracle_decode_file(path_particles, fused=1)
    plain = tl.oracle_decode_file(path_particles, fused=0)
    diff = fused != plain
    one = np.nonzero(diff.sum(axis=1) == 1)[0]
    r = np.float32(np.sqrt(np.float64(r2[shell])))
    out = []
    for i in one:
        d = int(np.nonzero(diff[i])[0][0])
        base = np.float32(fused[i, d] - r)
        cands = [base]
        up = dn = base
        for _ in range(64):
            up = np.nextafter(up, np.float32(np.inf))
            dn = np.nextafter(dn, np.float32(-np.inf))
            cands += [up, dn]
        for cd in cands:
            df = np.float32(fused[i, d] - cd)
            dp = np.float32(plain[i, d] - cd)
            if tl.shell_index(np.float32(df * df), r2) != tl.shell_index(np.float32(dp * dp), r2):
                c = fused[i].copy()
                c[d] = cd
                out.append(c)
                break
        if len(out) >= limit:
            break
    return np.array(out, np.float32)


def write_gz(src, dst):
    with open(src, "rb") as f, open(dst, "wb") as raw:
        with gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0, filename="") as g:
            g.write(f.read())


def run_ref(args, cwd, binary=REF):
    r = subprocess.run([binary] + args, cwd=cwd, capture_output=True, text=True, timeout=600)
    return r


def mask_stdout(s):
    return re.sub(r"(>>>> [a-z +]+ time: )\d+( milli seconds)", r"\1#\2", s)


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/astr_rcount is missing: make -f oracle/ref.mk first")
    rng = np.random.default_rng(20261203)
    cases = []
    tmp = tempfile.mkdtemp(prefix="sjcs_gold_")

    def case(name, particle_files, halo_file, radius, filter_file=None, extra=(), rt=False):
        args = ["--thread", "4", "--next-setup", "2", "--particle-files"] + \
               [os.path.join(GOLD, p) for p in particle_files] + \
               ["--halo-file", os.path.join(GOLD, halo_file), "--radius", radius]
        if filter_file:
            args += ["--filter-file", os.path.join(GOLD, filter_file)]
        args += list(extra)
        out = os.path.join(tmp, name + ".out")
        r = run_ref(args + ["--output-file", out], tmp)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: reference failed {r.returncode}: {r.stderr}")
        write_gz(out, os.path.join(GOLD, name + ".expected.gz"))
        stdout = mask_stdout(r.stdout).replace(GOLD + "/", "")
        entry = {"name": name, "particle_files": particle_files, "halo_file": halo_file,
                 "radius": radius, "filter_file": filter_file, "extra": list(extra),
                 "expected": name + ".expected.gz", "stdout_masked": stdout}
        if rt and os.path.exists(RT_REF):
            out_rt = os.path.join(tmp, name + ".rt.out")
            r2 = run_ref(args + ["--output-file", out_rt], tmp, RT_REF)
            if r2.returncode == 0:
                write_gz(out_rt, os.path.join(GOLD, name + ".rt_expected.gz"))
                entry["rt_expected"] = name + ".rt_expected.gz"
                entry["rt_stdout_masked"] = mask_stdout(r2.stdout).replace(GOLD + "/", "")
        cases.append(entry)
        print(f"{name}: ok", flush=True)

    # 1. the reference's own periodic-boundary data
    case("tdpb_8", ["tdpb_particles.dat"], "tdpb_halos.list", "0.001:5:8", rt=True)
    case("tdpb_4", ["tdpb_particles.dat"], "tdpb_halos.list", "0.5:2:4", rt=True)
    case("tdpb_swapped", ["tdpb_particles.dat"], "tdpb_halos.list", "5:0.001:8")

    # 2. uniform, config-1 shape scaled down: 10% of halos within 5 of a face
    p = rng.random((20000, 3), dtype=np.float32) * np.float32(1000.0)
    p[:8] = 1000.0
    p[8:16] = 0.0
    p[16:20] = [[0.0, 1000.0, 500.0], [1000.0, 0.0, 0.0], [999.99994, 0.0, 1000.0],
                [0.0, 0.0, 0.0]]
    write_particles(os.path.join(GOLD, "uni_particles.dat"), p)
    h = rng.random((2000, 3)) * 1000.0
    face = rng.random(2000) < 0.1
    ax = rng.integers(0, 3, 2000)
    lo = rng.random(2000) < 0.5
    for i in np.nonzero(face)[0]:
        h[i, ax[i]] = rng.uniform(0, 5) if lo[i] else rng.uniform(995, 1000)
    h[:4] = [[0, 0, 0], [1000, 1000, 1000], [0.0005, 999.9995, 500], [2.5, 2.5, 997.5]]
    write_halos(os.path.join(GOLD, "uni_halos.list"), h)
    case("uni_8", ["uni_particles.dat"], "uni_halos.list", "0.001:5:8", rt=True)
    case("uni_40", ["uni_particles.dat"], "uni_halos.list", "0.001:5:40")
    case("uni_dense_16", ["uni_particles.dat"], "uni_halos.list", "0.5:100:16", rt=True)
    case("uni_k1", ["uni_particles.dat"], "uni_halos.list", "3:7:1")
    case("uni_big_600", ["uni_particles.dat"], "uni_halos.list", "1:600:5")

    # 3. three particle files: counts accumulate over files
    for i in range(3):
        write_particles(os.path.join(GOLD, f"uni_part{i}.dat"),
                        p[i * 6000:(i + 1) * 6000])
    case("uni_3files", [f"uni_part{i}.dat" for i in range(3)], "uni_halos.list", "0.001:5:8",
         rt=True)

    # 4. ties and faces: dyadic grid points, so distances and box edges are
    #    exact; halos on grid points see particles exactly on their query
    #    box faces and at squared distances equal to shell thresholds
    g = np.arange(0, 24, dtype=np.float32) * np.float32(0.25)
    gx, gy, gz = np.meshgrid(g + 100.0, g + 200.0, g + 300.0, indexing="ij")
    grid = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1).astype(np.float32)
    extra_pts = np.array([[-3.0, 500, 500], [1003.0, 1.0, 2.0], [500, -0.5, 1000.5],
                          [1e30, 5, 5], [-1e30, 5, 5], [np.inf, 1.0, 1.0],
                          [-np.inf, 999, 999], [0.0, 0.0, 1000.0]], dtype=np.float32)
    write_particles(os.path.join(GOLD, "tie_particles.dat"), np.vstack([grid, extra_pts]))
    th = np.array([[102.0, 202.0, 302.0], [103.0, 203.0, 303.0], [101.25, 201.5, 302.75],
                   [100.0, 200.0, 300.0], [105.75, 205.75, 305.75], [0.0, 500.0, 500.0],
                   [999.0, 1.0, 2.0], [500.0, 0.0, 1000.0]])
    write_halos(os.path.join(GOLD, "tie_halos.list"), th)
    # radius 2^-2 .. 2^1: r2 = {1/16, .., 4} dyadic, r_box = 2 exactly
    case("tie_dyadic", ["tie_particles.dat"], "tie_halos.list", "0.25:2:4", rt=True)
    case("tie_dyadic_7", ["tie_particles.dat"], "tie_halos.list", "0.25:2:7")
    case("tie_mixed", ["tie_particles.dat"], "tie_halos.list", "0.1:1.75:9", rt=True)

    # 5. the reference's unreferenced irregular point set (negative and
    #    out-of-box coordinates), converted to the binary format
    if os.path.exists(RTREE_TEST):
        vals = []
        with open(RTREE_TEST) as f:
            for line in f:
                nums = re.findall(r"[-+]?\d+\.\d+(?:[eE][-+]?\d+)?", line)
                if len(nums) == 3:
                    vals.append([float(v) for v in nums])
        rp = np.array(vals, dtype=np.float32)
        write_particles(os.path.join(GOLD, "rtree_particles.dat"), rp)
        rh = np.vstack([rp[::97] + np.float32(0.01), [[0, 0, 0], [1, 1, 1], [999.5, 999.5, 0.5]]])
        write_halos(os.path.join(GOLD, "rtree_halos.list"), rh)
        case("rtree_4", ["rtree_particles.dat"], "rtree_halos.list", "0.05:4:12", rt=True)

    # 6. halo text edge cases: special values, odd number formats, and a last
    #    line without '\n' (the reference drops it)
    raw = ["900 0 1e+12 0 0 0 0 0 nan 500 500 0",
           "901 0 1e+12 0 0 0 0 0 inf 500 500 0",
           "902 0 1e+12 0 0 0 0 0 1e2 2e2 0x1.2cp+8 0",
           "903 0 1e+12 0 0 0 0 0 +250.5 -0 4.0e2 0",
           "904 0 1e+12 0 0 0 0 0 250.25abc 300 400 0",
           "18446744073709551615 0 1 0 0 0 0 0 250 250 250 0",
           "-1 0 1 0 0 0 0 0 251 251 251 0",
           "905 0 1e+12 0 0 0 0 0 260 260 260 0"]
    write_halos(os.path.join(GOLD, "fmt_halos.list"), h[:50], raw_lines=raw,
                trailing_newline=False)
    case("fmt_halos", ["uni_particles.dat"], "fmt_halos.list", "0.5:30:10", rt=True)

    # 7. --filter-file: unsorted ids with repeats and ids not present
    ids = list(range(0, 400, 3)) + [7, 7, 5000, 1]
    rng.shuffle(ids)
    with open(os.path.join(GOLD, "uni_filter.txt"), "w") as f:
        f.write("\n".join(str(i) for i in ids) + "\n")
    case("uni_filter", ["uni_particles.dat"], "uni_halos.list", "0.001:5:8",
         filter_file="uni_filter.txt")

    # 9. compressed (Gadget) particle files (--compressed).  A generator of
    #    its own, so the fixtures above stay as they were.
    crng = np.random.default_rng(20261204)
    cz = ["--compressed"]

    def uv_random(n, bits=16):
        return crng.integers(0, 1 << bits, (n, 3), dtype=np.uint16)

    # a whole 24^3 grid, spacing 1000/24 (inexact in FP32), displacements
    # +-30: particles below 0 and near 1000 on every axis
    write_compressed(os.path.join(GOLD, "gad_grid.gdt"), uv_random(24 ** 3), 24 ** 3)
    case("gad_grid", ["gad_grid.gdt"], "uni_halos.list", "0.5:20:12", extra=cz, rt=True)
    # one 32^3 snapshot in two files with their own displacement encodings
    uv2 = uv_random(32 ** 3, 12)
    write_compressed(os.path.join(GOLD, "gad_part0.gdt"), uv2[:16384], 32 ** 3,
                     disp_min=-20.0, disp_max=20.0, pos_bits=12, id_start=1)
    write_compressed(os.path.join(GOLD, "gad_part1.gdt"), uv2[16384:], 32 ** 3,
                     disp_min=-12.5, disp_max=7.75, pos_bits=12, id_start=16385)
    case("gad_2files", ["gad_part0.gdt", "gad_part1.gdt"], "uni_halos.list", "0.5:30:10",
         extra=cz, rt=True)
    # edges: id 0 (grid coordinate -1), ids past the grid (a second tile:
    # x beyond the box), extreme triplets; a half-size box with a negative
    # displacement span and 10-bit triplets
    uve = uv_random(8200)
    uve[:50] = 0
    uve[50:100] = 65535
    write_compressed(os.path.join(GOLD, "gad_edge0.gdt"), uve, 20 ** 3, disp_min=-50.0,
                     disp_max=50.0, id_start=0)
    write_compressed(os.path.join(GOLD, "gad_edge1.gdt"), uv_random(1000, 10), 10 ** 3,
                     box=500.0, disp_min=5.0, disp_max=-5.0, pos_bits=10, id_start=1)
    case("gad_edge", ["gad_edge0.gdt", "gad_edge1.gdt"], "uni_halos.list", "1:25:6", extra=cz)
    # a 1700^3 snapshot (npartTotal split over two words), one file of 6000
    # particles deep inside: 64-bit ids, inexact spacing; halos near them
    ng = 1700
    idb = 1 + ng * ng * 900 + ng * 850 + 100
    write_compressed(os.path.join(GOLD, "gad_big.gdt"), uv_random(6000), ng ** 3,
                     disp_min=-2.0, disp_max=2.0, id_start=idb)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import sjcs_testlib as tl
    pb = tl.oracle_decode_file(os.path.join(GOLD, "gad_big.gdt"), fused=1)
    hb = pb[crng.integers(0, pb.shape[0], 300)] + crng.uniform(-3, 3, (300, 3)).astype(np.float32)
    hb = np.vstack([hb, crng.random((100, 3)) * 1000.0])
    write_halos(os.path.join(GOLD, "gad_big_halos.list"), hb, fmt="%.9g")
    case("gad_big", ["gad_big.gdt"], "gad_big_halos.list", "0.05:3:10", extra=cz)
    # halos placed where the fused and the unfused decode disagree about a
    # shell: this case tells which of the two the reference computes
    write_compressed(os.path.join(GOLD, "gad_fma.gdt"), uv_random(48 * 48 * 2), 48 ** 3,
                     disp_min=-7.3, disp_max=7.3, id_start=1 + 48 * 48 * 20)
    _, r2f = tl.oracle_thresholds(0.5, 4.0, 8)
    hf = fma_sensitive_halos(os.path.join(GOLD, "gad_fma.gdt"), r2f, 5)
    if hf.shape[0] < 10:
        raise RuntimeError(f"gad_fma: only {hf.shape[0]} sensitive halos")
    write_halos(os.path.join(GOLD, "gad_fma_halos.list"), hf, fmt="%.9g")
    case("gad_fma", ["gad_fma.gdt"], "gad_fma_halos.list", "0.5:4:8", extra=cz)

    # 10. rt_rcount's leaf semantics: halos exactly on particles (a distance
    #     of 0 ends the scan of that particle's leaf), duplicated particles,
    #     clumps dense enough for deep trees, halos near faces.  Its own
    #     generator, so the fixtures above stay as they were.
    rrng = np.random.default_rng(20261205)
    cen = rrng.random((300, 3)) * 1000.0
    clump = cen[rrng.integers(0, 300, 15000)] + rrng.normal(0.0, 2.0, (15000, 3))
    rp = np.vstack([np.mod(clump, 1000.0), rrng.random((15000, 3)) * 1000.0])
    rp = rp.astype(np.float32)
    dup = rp[rrng.integers(0, rp.shape[0], 600)]
    rp = np.insert(rp, rrng.integers(0, rp.shape[0], 600), dup, axis=0)
    write_particles(os.path.join(GOLD, "rt_particles.dat"), rp)
    near = rp[rrng.integers(0, rp.shape[0], 200)].astype(np.float64)
    face = rrng.integers(0, 3, 200)
    for i in range(200):
        near[i, face[i]] = rrng.uniform(0, 3) if i % 2 else rrng.uniform(997, 1000)
    rh = np.vstack([rp[rrng.integers(0, rp.shape[0], 1500)].astype(np.float64),
                    cen[:400], rrng.random((700, 3)) * 1000.0, near])
    write_halos(os.path.join(GOLD, "rt_halos.list"), rh, fmt="%.9g")
    case("rt_onpart", ["rt_particles.dat"], "rt_halos.list", "0.01:3:10", rt=True)
    case("rt_onpart_wide", ["rt_particles.dat"], "rt_halos.list", "0.2:40:6", rt=True)
    # the same particles cut into two files: each file is its own tree
    write_particles(os.path.join(GOLD, "rt_part0.dat"), rp[:11000])
    write_particles(os.path.join(GOLD, "rt_part1.dat"), rp[11000:])
    case("rt_2files", ["rt_part0.dat", "rt_part1.dat"], "rt_halos.list", "0.01:3:10", rt=True)

    # 8. option errors: stderr and exit status (argv[0] shown as ARGV0)
    errors = []
    bad = [
        [],
        ["--thread", "4", "--next-setup", "2", "--particle-files", "x.dat", "--halo-file", "h"],
        ["--thread", "4", "--next-setup", "2", "--particle-files", "x.dat", "--halo-file", "h",
         "--radius", "5:5:8"],
        ["--thread", "0", "--next-setup", "2", "--particle-files", "x.dat", "--halo-file", "h",
         "--radius", "1:5:8"],
        ["--thread", "2", "--next-setup", "4", "--particle-files", "x.dat", "--halo-file", "h",
         "--radius", "1:5:8"],
        ["--thread", "4", "--next-setup", "2", "--halo-file", "h", "--radius", "1:5:8"],
        ["--thread", "4", "--next-setup", "2", "--particle-files", "x.dat", "--radius", "1:5:8"],
        ["--thread", "4", "--next-setup", "2", "--particle-files", "x.dat", "--halo-file", "h",
         "--radius", "1:5:8", "--backlog", "-1"],
        ["--thread", "4", "--next-setup", "3", "--particle-files", "x.dat", "--halo-file", "h",
         "--radius", "1:5:8", "--backlog", "1", "--first-setup", "2:1"],
        ["--thread", "4", "--next-setup", "2", "--particle-files", "--halo-file", "h",
         "--radius", "1:5"],
        ["--thread", "4", "--next-setup", "2", "--particle-files", "a", "b", "--halo-file", "h",
         "--radius", "-1:5:8", "--radius"],
    ]
    for i, args in enumerate(bad):
        r = run_ref(args, tmp)
        errors.append({"args": args, "returncode": r.returncode,
                       "stderr": r.stderr.replace(REF, "ARGV0"),
                       "stdout": r.stdout})
    with open(os.path.join(GOLD, "manifest.json"), "w") as f:
        json.dump({"cases": cases, "option_errors": errors}, f, indent=1)
    shutil.rmtree(tmp, ignore_errors=True)
    print("wrote", len(cases), "cases and", len(errors), "option-error cases")


if __name__ == "__main__":
    main()

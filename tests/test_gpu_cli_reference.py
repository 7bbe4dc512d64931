"""The drop-in CLI against the reference astr_rcount itself, on seeded
random inputs: both programs (the reference built from its sources into
oracle/_ref by oracle/ref.mk, run here as the checker) read the same
particle files (packed FP32, or compressed Gadget files,
src/Gadget/BufVector3D.c:111-136,285-308), the same halo text
(src/Gadget/BufHalo.c:76-119) and the same random argv -- `--radius
a:b:K`, `--thread`, `--next-setup`, `--backlog` (src/astr_rcount/main.c:
492-651) -- and must write byte-identical result files
(main.c:371-442), or, for an argv the reference refuses, exit with its
status and print its messages.  Skipped when the reference binary was not
built (it is built wherever the reference sources are present)."""
import os
import subprocess

import numpy as np
import pytest

import sjcs_testlib as t

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "astr_rcount")
RT_REF = os.path.join(ROOT, "oracle", "_ref", "rt_rcount")


def write_halo_text(path, ids, pts):
    with open(path, "w") as f:
        f.writelines(f"# header {i}\n" for i in range(16))
        f.writelines(f"{i} 1.5 0 0 0 0 0 0 {x:.9g} {y:.9g} {z:.9g} 0\n"
                     for i, (x, y, z) in zip(ids.tolist(), pts.tolist()))


def draw_run(built, seed, tmp_path):
    """Input files and argv of random case `seed` (without --output-file)."""
    rng = np.random.default_rng(90_000 + seed)
    nfiles = int(rng.integers(1, 4))
    compressed = rng.random() < 0.25
    files = []
    for i in range(nfiles):
        fn = str(tmp_path / f"part{i}.dat")
        if compressed:
            side = int(rng.integers(12, 40))
            disp = float(rng.uniform(0.5, 30.0))
            uv = rng.integers(0, 65536, (side ** 3, 3), dtype=np.uint16)
            built.write_compressed(fn, built.gadget_header(side ** 3, side ** 3, 1000.0, -disp,
                                                           disp, 16, 1), uv)
        else:
            n = int(10 ** rng.uniform(2.5, 4.6))
            p = rng.random((n, 3), dtype=np.float32) * np.float32(1000.0)
            if rng.random() < 0.3:  # a clump
                c = rng.random(3) * 1000.0
                k = n // 2
                p[:k] = np.mod(c + rng.standard_normal((k, 3)) * 2.0, 1000.0).astype(np.float32)
                p[:k] = np.minimum(p[:k], np.nextafter(np.float32(1000.0), np.float32(0)))
            p[:2] = 1000.0
            p[2:4] = 0.0
            p.tofile(fn)
        files.append(fn)
    nh = int(10 ** rng.uniform(1.0, 3.3))
    h = rng.random((nh, 3), dtype=np.float32) * np.float32(1000.0)
    m = np.nonzero(rng.random(nh) < 0.2)[0]
    h[m, rng.integers(0, 3, m.size)] = rng.choice([0.0, 1000.0, 1.0, 999.0], m.size)
    ids = np.sort(rng.choice(10 ** 9, nh, replace=False)).astype(np.uint64)
    hf = str(tmp_path / "halos.list")
    write_halo_text(hf, ids, h)
    rmax = float(10 ** rng.uniform(-0.5, 1.6))
    K = int(rng.integers(1, 49))
    radius = f"{rmax * float(10 ** rng.uniform(-3.5, -0.5)):.6g}:{rmax:.6g}:{K}"
    thr = str(int(rng.integers(1, 9)))
    args = ["--thread", thr, "--next-setup", str(int(rng.integers(1, 5))),
            "--particle-files", *files, "--halo-file", hf, "--radius", radius]
    if compressed:
        args.append("--compressed")
    backlog = int(rng.integers(0, 3))
    if backlog:
        args += ["--backlog", str(backlog)]
    print(f"case {seed}: files {nfiles} compressed {compressed} halos {nh} radius {radius} "
          f"threads {thr} backlog {backlog}", flush=True)
    return args


def compare_runs(ref, ours, ours_path, args, tmp_path):
    ref_out, our_out = tmp_path / "ref.txt", tmp_path / "ours.txt"
    r = subprocess.run([ref] + args + ["--output-file", str(ref_out)], capture_output=True,
                       text=True, timeout=120)
    o = subprocess.run([ours] + args + ["--output-file", str(our_out)],
                       capture_output=True, text=True, timeout=120)
    assert o.returncode == r.returncode, o.stderr[-500:]
    if r.returncode != 0:
        # an argv the reference refuses (e.g. --backlog with --thread not
        # above the index construction threads, main.c:576-589): the same
        # messages, the same usage text
        assert o.stderr.replace(ours_path, "ARGV0") == r.stderr.replace(ref, "ARGV0")
        assert o.stdout == r.stdout
        return
    a, b = ref_out.read_bytes(), our_out.read_bytes()
    assert len(a) > 0 and a == b, f"result files differ ({len(a)} vs {len(b)} bytes)"


@pytest.mark.skipif(not os.path.exists(REF), reason="reference binary not built (oracle/ref.mk)")
@pytest.mark.parametrize("seed", range(48))
def test_cli_equals_reference_binary(built, seed, tmp_path):
    args = draw_run(built, seed, tmp_path)
    compare_runs(REF, built.CLI_PATH, built.CLI_PATH, args, tmp_path)


@pytest.mark.skipif(not os.path.exists(RT_REF), reason="reference binary not built (oracle/ref.mk)")
@pytest.mark.parametrize("seed", range(24))
def test_rt_cli_equals_reference_binary(built, seed, tmp_path):
    """The CLI run as rt_rcount (a link of that name) against the
    reference rt_rcount (src/rt_rcount/main.c) on the same random cases:
    its R-tree counts, byte for byte."""
    link = tmp_path / "rt_rcount"
    os.symlink(built.CLI_PATH, link)
    args = draw_run(built, 1000 + seed, tmp_path)
    compare_runs(RT_REF, str(link), str(link), args, tmp_path)

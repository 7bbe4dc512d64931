#!/usr/bin/env python3
"""bench.py -- shell pair-tests/s of the SJCS hot path on one MI355X.

Workload (default): the reference README's search shape, one chunk --
10M halos x 100M particles x 40 shells (``--radius 0.001:5:40``), halo and
particle coordinates uniform FP32 in [0, 1000)^3, synthetic (seeded numpy).

A step is one pass of the hot path over the resident particle chunk: the
cell sort (three MSD passes by cz, cy, cx, writing the cell table), and the
count of every halo into the u64 counts (stored, as for the first particle
file).  Inputs are in HBM before the timed region starts; the halos are
sorted once beforehand, as the reference reads its halo file once before its
chunk loop.

Config c2z is the same search with the particles given as a compressed
(Gadget) snapshot: a step then starts from the device-resident uint16
triplets, which the cell sort's first pass decodes as it reads them.

``value`` = W / (seconds per step), W = exact number of in-box
(halo image, particle) pair tests of the chunk -- the reference's
CalcDistance calls -- counted by an untimed pass.

Prints one JSON line.  ``--gpus`` above 1 is refused: the build targets one
GPU.  ``--calibrate-copy MIB`` instead runs one plain device-to-device copy
of MIB MiB (torch), the reference point for the FETCH_SIZE / WRITE_SIZE
counters behind ``roofline.traffic``.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import spatialjoincountovershells_amd as sj  # noqa: E402
from spatialjoincountovershells_amd import synth  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "astr_rcount")
CLI_BIN = os.path.join(ROOT, "spatialjoincountovershells_amd", "bin", "sjcs_rcount")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3

CONFIGS = {
    "c2": dict(workload="uniform 1e8 particles x 1e7 halos, --radius 0.001:5:40, one chunk",
               np=100_000_000, nh=10_000_000, radius=(0.001, 5.0, 40), chunks=1),
    "c1": dict(workload="uniform 1e5 particles x 1e4 halos, --radius 0.001:5:8, one chunk",
               np=100_000, nh=10_000, radius=(0.001, 5.0, 8), chunks=1),
    "c4": dict(workload="clustered 2e8 particles (70% in 200k Gaussian clumps, sigma 0.3-2, "
                        "power-law sizes) x 5e6 halos (80% at clump centres), "
                        "--radius 0.01:20:64, one chunk",
               np=200_000_000, nh=5_000_000, radius=(0.01, 20.0, 64), chunks=1,
               clustered=True, autotune=True, compare_plain=True),
    "c5": dict(workload="serving: 3e8 uniform particles resident; requests of floor(10^U[0,5]) "
                        "uniform halos, K ~ U{8..64}, r_max ~ U[1,10], r_min = r_max/5000; "
                        "Poisson arrivals",
               np=300_000_000, nh=0, radius=(0.002, 10.0, 64), chunks=1, serving=True),
    "c3": dict(workload="uniform 3 x 1e8 particles (3 files) x 1e7 halos, --radius 0.001:5:40, "
                        "u64 counts accumulated on the device",
               np=100_000_000, nh=10_000_000, radius=(0.001, 5.0, 40), chunks=3, graph=True),
    "c2z": dict(workload="compressed Gadget snapshot: 464^3 grid (99.9M particles, uint16 "
                         "displacements over +-20) x 1e7 uniform halos, --radius 0.001:5:40, "
                         "decoded on the device",
                np=464 ** 3, nh=10_000_000, radius=(0.001, 5.0, 40), chunks=1, compressed=464),
    "rt": dict(workload="rt_rcount mode: uniform 1e7 particles in one R-tree (built on the host "
                        "in insertion order) x 1e6 uniform halos (a tenth exactly on particles), "
                        "--radius 0.001:5:40",
               np=10_000_000, nh=1_000_000, radius=(0.001, 5.0, 40), chunks=1, rtree=True),
}


METRICS = {
    "c1": "shell pair-tests/s (10k halos x 100k particles x 8 shells, 1 GPU)",
    "c2": "shell pair-tests/s (10M halos x 100M particles x 40 shells, 1 GPU)",
    "c3": "shell pair-tests/s (10M halos x 3 x 100M particles x 40 shells, 1 GPU)",
    "c4": "shell pair-tests/s (5M clustered halos x 200M clustered particles x 64 shells, 1 GPU)",
    "c5": "served requests/s (3e8-particle resident index, batches of 1-1e5 halos, own shells)",
    "c2z": "shell pair-tests/s (10M halos x 99.9M compressed particles x 40 shells, 1 GPU)",
    "rt": "rt_rcount halo queries/s (1M halos x 10M-particle R-tree x 40 shells, 1 GPU)",
}
RT_REF_BIN = os.path.join(ROOT, "oracle", "_ref", "rt_rcount")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def csrc_digest():
    """sha256 of the HIP sources (csrc/): counter figures in profiles/*.json
    carry the digest of the sources they were measured on."""
    import hashlib
    d = os.path.join(ROOT, "spatialjoincountovershells_amd", "csrc")
    h = hashlib.sha256()
    for name in sorted(os.listdir(d)):
        with open(os.path.join(d, name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()


def measured_profile(name):
    """profiles/<name> when it was measured on the current sources, else
    None (a stale counter figure is not reported)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        js = json.load(f)
    if js.get("csrc_sha256") != csrc_digest():
        log(f"[bench] {name}: measured on other sources, not reported")
        return None
    return js


def host_threads():
    """Host cores this process may run on (the CPU baseline's thread count)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def make_inputs(cfg, seed):
    """Particle files (one array each, distinct seeds) and the halos."""
    rng = np.random.default_rng(seed)
    if cfg.get("compressed"):
        part = synth.gadget_snapshot(rng, cfg["compressed"])
        return [part], synth.uniform_points(rng, cfg["nh"])
    if cfg.get("clustered"):
        parts, halos = synth.clustered(rng, cfg["np"], cfg["nh"])
        return [parts], halos
    parts = [synth.uniform_points(rng, cfg["np"])]
    halos = synth.uniform_points(rng, cfg["nh"])
    for c in range(1, cfg["chunks"]):
        parts.append(synth.uniform_points(np.random.default_rng(seed + 1000 * c), cfg["np"]))
    return parts, halos


def write_halo_file(path, halos):
    """The halo text format the reference reads (src/Gadget/BufHalo.c:76-119):
    16 header lines, then id and 7 other columns with x, y, z in columns 9-11;
    %.9g gives back the same FP32 values through strtof."""
    with open(path, "w") as f:
        for i in range(16):
            f.write(f"# header {i}\n")
        step = 1 << 20
        for a in range(0, halos.shape[0], step):
            blk = halos[a:a + step].tolist()
            f.write("".join(f"{a + i} 0 0 0 0 0 0 0 {x:.9g} {y:.9g} {z:.9g} 0\n"
                            for i, (x, y, z) in enumerate(blk)))


def timer_lines(text):
    return {m.group(1): int(m.group(2))
            for m in re.finditer(r">>>> ([a-z +]+) time: (\d+) milli seconds", text)}


def is_compressed(part):
    return isinstance(part, tuple)


def write_particle_files(tmp, parts):
    """One file per chunk: packed FP32 xyz, or a compressed (Gadget) file for
    a (header, triplets) chunk.  Returns the paths and the CLI flag."""
    pfs = []
    for c, p in enumerate(parts):
        if is_compressed(p):
            pfs.append(os.path.join(tmp, f"part{c}.gdt"))
            sj.write_compressed(pfs[-1], *p)
        else:
            pfs.append(os.path.join(tmp, f"part{c}.dat"))
            p.tofile(pfs[-1])
    flag = ["--compressed"] if parts and is_compressed(parts[0]) else []
    return pfs, flag


def chunk_size(part):
    return part[1].shape[0] if is_compressed(part) else part.shape[0]


def add_device_chunk(eng, d, part):
    """Queues one device-resident chunk: decoded points, or triplets decoded
    on the device."""
    if is_compressed(part):
        eng.add_compressed_particles_device(d, part[1].shape[0], sj.gadget_decode_params(part[0]))
    else:
        eng.add_particles_device(d, part.shape[0])


def wall_legs(parts, halos, radius, threads, with_reference):
    """End-to-end process wall time of the drop-in CLI (and, if asked, of the
    reference binary built from its sources) on files written here: reading
    the halo text and the particle files, the count, formatting the result
    (to /dev/null: disk speed is not measured)."""
    tmp = tempfile.mkdtemp(prefix="sjcs_wall_")
    try:
        pfs, flag = write_particle_files(tmp, parts)
        hf = os.path.join(tmp, "halo.list")
        write_halo_file(hf, halos)
        mn, mx, k = radius
        args = ["--thread", str(threads), "--next-setup", str(threads), "--particle-files", *pfs,
                *flag, "--halo-file", hf, "--radius", f"{mn}:{mx}:{k}",
                "--output-file", os.devnull]
        res = {}
        legs = [("sjcs_rcount", CLI_BIN, [])]
        if len(parts) > 1:
            # files held on the device ahead of the count (the reference's
            # background index builds)
            legs.append(("sjcs_rcount --backlog 2", CLI_BIN, ["--backlog", "2"]))
        if with_reference and os.path.exists(REF_BIN):
            legs.append(("reference astr_rcount", REF_BIN, ["--backlog", "0"]))
        for name, exe, extra in legs:
            t0 = time.time()
            out = subprocess.run([exe] + extra + args, capture_output=True, text=True, timeout=900)
            wall = time.time() - t0
            if out.returncode != 0:
                log(f"{name} failed rc={out.returncode}: {out.stderr[-300:]}")
                continue
            res[name] = {"wall_s": round(wall, 3), "phases_ms": timer_lines(out.stdout),
                         "threads": threads}
        return res
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_baseline(parts, halos, radius, eng, d_parts, nsample, threads):
    """Times the reference's astr_rcount (built from its sources into
    oracle/_ref) on the same particle files and a halo sample; returns the
    cpu_baseline object, or None when the binary is absent."""
    if not os.path.exists(REF_BIN):
        return None
    tmp = tempfile.mkdtemp(prefix="sjcs_cpu_")
    try:
        pfs, flag = write_particle_files(tmp, parts)
        hf = os.path.join(tmp, "halo.list")
        # a random sample (the clustered set lists its clump-centre halos first)
        pick = np.sort(np.random.default_rng(7).choice(halos.shape[0], nsample, replace=False))
        hs = halos[pick]
        with open(hf, "w") as f:
            for i in range(16):
                f.write(f"# dummy header {i}\n")
            # %.9g round-trips FP32 exactly, so strtof gives back these values
            lines = [f"{i} 0 0 0 0 0 0 0 {x:.9g} {y:.9g} {z:.9g} 0\n"
                     for i, (x, y, z) in enumerate(hs.tolist())]
            f.writelines(lines)
        mn, mx, k = radius
        cmd = [REF_BIN, "--thread", str(threads), "--next-setup", str(threads),
               "--backlog", "0", "--particle-files", *pfs, *flag, "--halo-file", hf,
               "--radius", f"{mn}:{mx}:{k}", "--output-file", os.devnull]
        t0 = time.time()
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        wall = time.time() - t0
        m = re.findall(r">>>> searching time: (\d+) milli seconds", out.stdout)
        if out.returncode != 0 or not m:
            log(f"cpu baseline failed rc={out.returncode}: {out.stderr[-300:]}")
            return None
        search_s = sum(int(v) for v in m) / 1000.0
        # exact W of the sample, counted on the GPU with the same floats
        eng.set_halos(hs)
        w = 0
        for d, p in zip(d_parts, parts):
            add_device_chunk(eng, d, p)
            w += eng.last_chunk_pair_tests()
        eng.set_halos(halos)
        npart = sum(chunk_size(p) for p in parts)
        return {"value": w / search_s, "unit": "pair-tests/s", "cores": threads,
                "kind": "reference",
                "sample": f"astr_rcount --thread {threads}: all {npart} particles "
                          f"({len(parts)} file(s)) x {nsample} halos (random sample of the "
                          f"same set), {mn}:{mx}:{k}; W={w} in {search_s:.3f} s of "
                          f"'searching time' (process wall {wall:.1f} s)"}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def serve_bench(args, cfg):
    """Config 5: a resident index and a stream of requests through the
    serving loop.  First the capacity (all requests submitted at once), then
    Poisson arrivals at `load` times that rate: latency p50/p99 from the
    server's own submit and completion times."""
    rng = np.random.default_rng(args.seed)
    log(f"[bench] generating {cfg['np']} particles")
    parts = synth.uniform_points(rng, cfg["np"])
    reqs = []
    for _ in range(args.requests):
        n = int(10 ** rng.uniform(0.0, 5.0))
        k = int(rng.integers(8, 65))
        rmax = float(rng.uniform(1.0, 10.0))
        _, r2 = sj.shell_thresholds(rmax / 5000.0, rmax, k)
        reqs.append((synth.uniform_points(rng, n), r2))
    halos = sum(h.shape[0] for h, _ in reqs)
    eng = sj.Engine(0)
    ix = sj.ArraySTR(eng, parts.shape[0])
    t0 = time.perf_counter()
    ix.insert(parts)
    ix.make_tree()
    build_s = time.perf_counter() - t0
    del parts
    sv = sj.Server(ix)
    # warm-up: the slots' buffers grow to the largest request
    for h, r2 in sorted(reqs, key=lambda q: -q[0].shape[0])[:16]:
        tk = sv.submit(h, r2)
        sv.wait(tk)
        sv.release(tk)
    # capacity: every request submitted at once, twice.  The first pass's
    # output arrays are fresh host memory, each page faulted in as the
    # counts land (glibc maps large blocks anew); the second pass reuses
    # what the first returned, as a caller recycling its result buffers
    # does, and is the figure reported.
    cap_runs = []
    for rep in range(2):
        log(f"[bench] capacity run {rep + 1}")
        t0 = time.perf_counter()
        tickets = [sv.submit(h, r2) for h, r2 in reqs]
        for tk in tickets:
            sv.wait(tk)
        cap_runs.append(time.perf_counter() - t0)
        for tk in tickets:
            sv.release(tk)
    cap_s = cap_runs[-1]
    capacity = len(reqs) / cap_s
    rate = args.load * capacity
    log(f"[bench] capacity {capacity:.1f} req/s; Poisson arrivals at {rate:.1f} req/s")
    gaps = rng.exponential(1.0 / rate, len(reqs))
    t0 = time.perf_counter()
    due = 0.0
    tickets = []
    for (h, r2), g in zip(reqs, gaps):
        due += g
        wait = due - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        tickets.append(sv.submit(h, r2))
    lat = []
    for tk in tickets:
        sv.wait(tk)
        sub, adm, done = sv.times(tk)
        lat.append(done - sub)
    run_s = time.perf_counter() - t0
    lat = np.array(lat) * 1e3
    sv.close()
    ix.close()
    eng.close()
    out = {
        "metric": METRICS[args.config],
        "value": capacity,
        "unit": "requests/s",
        "n_gpus": 1,
        "steps": len(reqs),
        "warmup": 16,
        "ms_per_step": 1e3 / capacity,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "particles": cfg["np"], "requests": len(reqs),
                   "halos": halos, "slots": 8, "parallelism": "single GPU"},
        "capacity": {"requests_per_s": capacity, "halos_per_s": halos / cap_s,
                     "wall_s": round(cap_s, 3),
                     "first_pass_requests_per_s": len(reqs) / cap_runs[0],
                     "note": "second of two passes; the first writes into fresh host memory"},
        "poisson": {"rate_per_s": rate, "load": args.load, "wall_s": round(run_s, 3),
                    "latency_ms": {"p50": float(np.percentile(lat, 50)),
                                   "p99": float(np.percentile(lat, 99)),
                                   "max": float(lat.max())}},
        "index_build_s": round(build_s, 3),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def rt_bench(args, cfg):
    """rt_rcount mode: one R-tree over the particles, built on the host point
    by point as rt_rcount builds it, then every halo counted against it on
    the GPU (sjcs_rtree_count_device: centres and u64 counts resident in
    HBM).  A step is one count of all halos; the tree's upload is the first
    warm-up step.  The CPU baseline is the reference rt_rcount (built from
    its sources) on a halo sample of the same set, whose result rows must
    equal the GPU's."""
    rng = np.random.default_rng(args.seed)
    mn, mx, k = cfg["radius"]
    _, r2 = sj.shell_thresholds(mn, mx, k)
    log(f"[bench] generating {cfg['workload']}")
    parts = synth.uniform_points(rng, cfg["np"])
    halos = synth.uniform_points(rng, cfg["nh"])
    on = rng.choice(cfg["nh"], cfg["nh"] // 10, replace=False)
    halos[on] = parts[rng.integers(0, cfg["np"], on.size)]
    nh = halos.shape[0]
    eng = sj.Engine(0)
    rt = sj.RTree(eng)
    t0 = time.perf_counter()
    rt.insert(parts)
    build_s = time.perf_counter() - t0
    shape = rt.shape()
    log(f"[bench] host R-tree build {build_s:.2f} s: {shape}")
    d_h = eng.device_array(halos)
    d_c = eng.device_array(np.zeros((nh, k), np.uint64))
    t0 = time.perf_counter()
    rt.multi_range_count_device(d_h, nh, r2, d_c)
    first_s = time.perf_counter() - t0
    for _ in range(max(args.warmup - 1, 0)):
        rt.multi_range_count_device(d_h, nh, r2, d_c)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        rt.multi_range_count_device(d_h, nh, r2, d_c)  # returns when the counts are added
        times.append(time.perf_counter() - t0)
    ms = 1e3 * sum(times) / len(times)
    eng.device_free(d_c)
    counts = rt.multi_range_count(halos, r2)
    eng.device_free(d_h)
    cpu = None
    if not args.no_cpu_baseline and os.path.exists(RT_REF_BIN):
        log("[bench] cpu baseline (reference rt_rcount on a halo sample)")
        cpu = rt_cpu_baseline(parts, halos, counts, cfg["radius"], args)
    rt.close()
    eng.close()
    out = {
        "metric": METRICS[args.config],
        "value": nh / (ms / 1e3),
        "unit": "halos/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "particles": cfg["np"], "halos": nh,
                   "shells": k, "radius": f"{mn}:{mx}:{k}", "pairs_counted": int(counts.sum()),
                   "parallelism": "single GPU"},
        "tree": dict(shape, host_build_s=round(build_s, 3),
                     first_count_with_upload_s=round(first_s, 3)),
        "roofline": None,
        "roofline_note": "a walk of dependent node loads: latency-bound, no byte or flop "
                         "roofline applies",
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def rt_cpu_baseline(parts, halos, counts, radius, args):
    tmp = tempfile.mkdtemp(prefix="sjcs_rt_")
    try:
        pf = os.path.join(tmp, "part0.dat")
        parts.tofile(pf)
        nsample = min(args.cpu_halos, 20_000, halos.shape[0])
        pick = np.sort(np.random.default_rng(7).choice(halos.shape[0], nsample, replace=False))
        hf = os.path.join(tmp, "halo.list")
        write_halo_file(hf, halos[pick])
        out = os.path.join(tmp, "out.txt")
        mn, mx, k = radius
        threads = args.cpu_threads
        cmd = [RT_REF_BIN, "--thread", str(threads), "--next-setup", str(threads),
               "--backlog", "0", "--particle-files", pf, "--halo-file", hf,
               "--radius", f"{mn}:{mx}:{k}", "--output-file", out]
        t0 = time.time()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        wall = time.time() - t0
        m = re.findall(r">>>> searching time: (\d+) milli seconds", r.stdout)
        b = re.findall(r">>>> index construction time: (\d+) milli seconds", r.stdout)
        if r.returncode != 0 or not m:
            log(f"rt cpu baseline failed rc={r.returncode}: {r.stderr[-300:]}")
            return None
        with open(out, "rb") as f:
            text = f.read().decode()
        rows = [int(v) for v in re.findall(r": (\d+)\n", text.split("\nTotal\n")[0])]
        same = np.array_equal(np.array(rows, np.uint64).reshape(nsample, k), counts[pick])
        if not same:
            raise SystemExit("bench.py: rt counts differ from the reference rt_rcount's")
        search_s = max(int(m[0]), 1) / 1000.0
        return {"value": nsample / search_s, "unit": "halos/s", "cores": threads,
                "kind": "reference",
                "sample": f"rt_rcount --thread {threads}: all {parts.shape[0]} particles x "
                          f"{nsample} halos (random sample of the same set), {mn}:{mx}:{k}; "
                          f"{search_s:.3f} s of 'searching time', its single-threaded index "
                          f"construction {int(b[0]) / 1000.0 if b else float('nan'):.1f} s "
                          f"(process wall {wall:.1f} s); its rows equal the GPU's"}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cell-size-milli", type=int, default=0)
    ap.add_argument("--lanes-per-halo", type=int, default=0)
    ap.add_argument("--plain", action="store_true",
                    help="the plain count: one halo per wavefront, no halo tiles, no autotuning")
    ap.add_argument("--autotune", action="store_true",
                    help="run the engine's autotuner on the first particle file before timing")
    ap.add_argument("--graph", action="store_true",
                    help="replay each chunk's sort + count from a captured HIP graph")
    ap.add_argument("--overlap", action="store_true",
                    help="sort the next chunk while this one is counted")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-halos", type=int, default=1_000_000)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline and the CLI (0 = every core "
                         "this process may use)")
    ap.add_argument("--plain-steps", type=int, default=2,
                    help="c4: timed steps of the plain kernel (one halo per wavefront, no "
                         "tiles, default cells) after the tuned ones; 0 = skip")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB")
    ap.add_argument("--requests", type=int, default=2000, help="serving config: requests")
    ap.add_argument("--wall", action="store_true",
                    help="also time the drop-in CLI end to end on files of this workload "
                         "(default for c2)")
    ap.add_argument("--no-wall", action="store_true")
    ap.add_argument("--wall-reference", action="store_true",
                    help="with --wall: time the reference binary on the same files too")
    ap.add_argument("--load", type=float, default=0.7,
                    help="serving config: arrival rate as a fraction of the measured capacity")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(args.calibrate_copy * 262144, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": a.numel() * 4}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus}: this build runs on one GPU only "
              "(multi-GPU sharding is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)

    if args.cpu_threads <= 0:
        args.cpu_threads = host_threads()
    cfg = CONFIGS[args.config]
    if cfg.get("serving"):
        return serve_bench(args, cfg)
    if cfg.get("rtree"):
        return rt_bench(args, cfg)
    mn, mx, k = cfg["radius"]
    _, r2 = sj.shell_thresholds(mn, mx, k)
    log(f"[bench] generating {cfg['workload']}")
    parts, halos = make_inputs(cfg, args.seed)

    eng = sj.Engine(0)
    if args.cell_size_milli:
        eng.set_option("cell_size_milli", args.cell_size_milli)
    if args.lanes_per_halo:
        eng.set_option("lanes_per_halo", args.lanes_per_halo)
    if args.plain:
        eng.set_option("kernel", 1)
        eng.set_option("tiles", 0)
    if args.overlap:
        eng.set_option("overlap", 1)
    if args.graph or cfg.get("graph"):
        eng.set_option("graph", 1)
    eng.set_shells(r2)
    # the first call also loads the sort's kernels into the process; the
    # second is timed
    eng.set_halos(halos)
    t_h = time.perf_counter()
    eng.set_halos(halos)
    halo_wall_ms = (time.perf_counter() - t_h) * 1e3
    halo_sort_ms = eng.phase_timings()["halo_sort"]
    d_parts = [eng.device_array(p[1] if is_compressed(p) else p) for p in parts]
    n = cfg["np"]

    def step():
        eng.reset_counts()
        for d, p in zip(d_parts, parts):
            add_device_chunk(eng, d, p)

    tuned = None
    if args.autotune and cfg.get("compressed"):
        sys.exit("bench.py: --autotune takes decoded particles")
    if args.autotune or (cfg.get("autotune") and not args.plain):
        t_at = time.perf_counter()
        best = eng.autotune(d_parts[0], n)
        tuned = {"wall_s": round(time.perf_counter() - t_at, 3), "best_ms": round(best, 3),
                 "chosen": {k: eng.get_option(k) for k in
                            ("cell_size_milli", "x_split", "lanes_per_halo", "segment", "tiles",
                             "tile_level", "tile_halos")}}
        log(f"[bench] autotune: {tuned}")

    # exact W of one step: an untimed pair-test pass per chunk
    w_chunks = []
    for d, p in zip(d_parts, parts):
        eng.reset_counts()
        add_device_chunk(eng, d, p)
        w_chunks.append(eng.last_chunk_pair_tests())
    w = sum(w_chunks)

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
        eng.sync()

    log(f"[bench] timing {args.steps} steps")
    phase = np.zeros((args.steps, 3))
    eng.sync()
    t0 = time.perf_counter()
    for s_ in range(args.steps):
        step()
        phase[s_] = eng.last_chunk_timings()  # waits for the step (last chunk's phases)
    eng.sync()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed * 1000.0 / args.steps

    t_r = time.perf_counter()
    counted = int(eng.counts().sum())
    counts_wall_ms = (time.perf_counter() - t_r) * 1e3
    readout_ms = eng.phase_timings()["readout"]
    build_ms = float(phase[:, 1].mean())
    count_ms = float(phase[:, 2].mean())
    graph = bool(args.graph or cfg.get("graph"))
    log(f"[bench] W={w} counted={counted} step={ms_step:.3f} ms build={build_ms:.3f} "
        f"count={count_ms:.3f} ms (last chunk)")

    plain = None
    if cfg.get("compare_plain") and not args.plain and args.plain_steps > 0:
        # the plain one-halo-per-wavefront kernel on the same inputs, same
        # run: default cells, no tiles
        tuned_opts = {k_: eng.get_option(k_) for k_ in
                      ("cell_size_milli", "x_split", "lanes_per_halo", "segment", "tiles",
                       "tile_level", "tile_halos")}
        for k_ in ("cell_size_milli", "x_split", "lanes_per_halo", "segment", "tile_level",
                   "tile_halos"):
            eng.set_option(k_, 0)
        eng.set_option("tiles", 0)
        eng.set_option("kernel", 1)
        log(f"[bench] plain kernel: {args.plain_steps} steps")
        step()
        eng.sync()
        pph = np.zeros((args.plain_steps, 3))
        t0 = time.perf_counter()
        for s_ in range(args.plain_steps):
            step()
            pph[s_] = eng.last_chunk_timings()
        eng.sync()
        p_ms = (time.perf_counter() - t0) * 1e3 / args.plain_steps
        p_counted = int(eng.counts().sum())
        plain = {"ms_per_step": round(p_ms, 3), "count_ms": round(float(pph[:, 2].mean()), 3),
                 "index_build_ms": round(float(pph[:, 1].mean()), 3),
                 "pair_tests_per_s": w / (p_ms * 1e-3),
                 "same_counts_total": p_counted == counted,
                 "step_speedup_tuned_vs_plain": round(p_ms / ms_step, 3),
                 "count_speedup_tuned_vs_plain": round(float(pph[:, 2].mean()) / count_ms, 3),
                 "target": ">= 2x (batched persistent kernel vs plain)"}
        eng.set_option("kernel", 0)
        for k_, v_ in tuned_opts.items():
            eng.set_option(k_, v_)
        log(f"[bench] plain: {plain}")

    nh = halos.shape[0]
    # algorithmic bytes of one count launch: particles read once, halo SoA
    # x, y, z read once, u32 running counts written once (read too when a
    # later chunk adds into them)
    count_bytes = 12 * n + 12 * nh + 4 * k * nh * (2 if cfg["chunks"] > 1 else 1)
    count_gbs = count_bytes / (count_ms * 1e-3) / 1e9
    count_tflops = 8 * w_chunks[-1] / (count_ms * 1e-3) / 1e12
    # index build: packed xyz (12 B) or compressed triplets (6 B) read once,
    # sorted SoA xyz written once
    build_bytes = (6 if cfg.get("compressed") else 12) * n + 12 * n
    build_gbs = build_bytes / (build_ms * 1e-3) / 1e9
    dominant = "count" if count_ms >= build_ms else "index_build"
    traffic = None
    tj = measured_profile("traffic.json")
    if tj is not None:
        traffic = tj.get(args.config, {}).get(dominant)
    if dominant == "count":
        roof = {"bound": "hbm", "achieved": round(count_gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(count_gbs / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "kernel": "count_group_kernel" + (" (graph replay: timed together with the "
                                                  "cell sort)" if graph else ""),
                "bytes_per_launch": count_bytes}
    else:
        roof = {"bound": "hbm", "achieved": round(build_gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(build_gbs / HBM_PEAK_GBS, 4),
                "traffic": traffic, "kernel": "cell sort (cs_* kernels)",
                "bytes_per_launch": build_bytes}

    # measured VALU issue of the count kernel (rocprofv3 counters, profiles/)
    vj = measured_profile("valu.json")
    valu = vj.get(args.config, {}).get("count") if vj is not None else None

    cpu = None
    if not args.no_cpu_baseline and args.config in ("c2", "c3", "c4", "c2z"):
        log("[bench] cpu baseline (reference astr_rcount on a halo sample)")
        nsample = args.cpu_halos if args.config != "c4" else min(args.cpu_halos, nh // 100)
        cpu = cpu_baseline(parts, halos, cfg["radius"], eng, d_parts,
                           min(nsample, nh), args.cpu_threads)

    for d in d_parts:
        eng.device_free(d)
    eng.close()
    wall = None
    if (args.wall or args.config == "c2") and not args.no_wall:
        log("[bench] end-to-end wall time of the CLI")
        wall = wall_legs(parts, halos, cfg["radius"], args.cpu_threads, args.wall_reference)
    value = w / (ms_step * 1e-3)
    out = {
        "metric": METRICS[args.config],
        "value": value,
        "unit": "pair-tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "particles": n * cfg["chunks"],
                   "chunks": cfg["chunks"], "halos": nh, "shells": k,
                   "radius": f"{mn}:{mx}:{k}", "pair_tests_per_step": w,
                   "pairs_counted": counted, "parallelism": "single GPU"},
        "phases_ms": {
            **({"index_build+count (graph replay)": round(count_ms, 4)} if graph else
               {"index_build": round(build_ms, 4), "count": round(count_ms, 4)}),
            "of": "the last chunk of a step" if cfg["chunks"] > 1 else "the chunk",
            "outside_the_step": {
                "halo_sort_device": round(halo_sort_ms, 4),
                "halo_upload_and_sort_wall": round(halo_wall_ms, 3),
                "readout_device": round(readout_ms, 4),
                "readout_and_download_wall": round(counts_wall_ms, 3),
                "note": "once per halo set / per result, not per step: the halos are "
                        "sorted before the chunk loop, the counts put back in input "
                        "order and copied to the host after it"}},
        "count_kernel": {"of": "the last chunk of a step" + (" (graph replay: its sort "
                                                              "included)" if graph else ""),
                         "pair_tests_per_s": w_chunks[-1] / (count_ms * 1e-3),
                         "valu_tflops_8_per_test": round(count_tflops, 3),
                         "valu_frac_of_fp32_peak": round(count_tflops / FP32_PEAK_TFLOPS, 4),
                         "hbm_gbs_algorithmic": round(count_gbs, 1),
                         "valu_busy_measured": valu and valu.get("valu_busy"),
                         "valu_lane_slots_per_pair_test":
                             valu and valu.get("valu_lane_slots_per_pair_test")},
        "autotune": tuned,
        "plain_kernel": plain,
        "wall": wall,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark: batched global gap-affine alignment (scores + Ties.LEFT CIGARs) on one MI355X.

Workload (one step): 100,000 pairs, 250 bp uniform ACGT references, targets mutated at 5% a base (60%
substitutions, 20% insertions, 20% deletions, indel lengths Geometric(0.7)), costs mismatch 4, gap-open 6,
gap-extend 2, every pair's cost and CIGAR. The pairs are packed and uploaded once (dinara_batch_new); a step
is one dinara_batch_run, the whole device path with its inputs resident in HBM: the banded sweeps, the band
certificate and its retries, the walk back and the CIGAR text.

The same run times the restated CPU batch path (oracle/baseline.cpp, compiled -march=native here) on all
host cores over the same pairs and checks every GPU result against it byte for byte.

Prints one JSON line. `--gpus` above 1 exits: this build runs on one GPU.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _path in (ROOT, os.path.join(ROOT, "dinara-align_amd")):
    if _path not in sys.path:
        sys.path.insert(0, _path)

from dinara_align import _lib, synthetic  # noqa: E402

MISMATCH, OPEN, EXTEND = 4, 6, 2
VALU_PEAK_TOPS = 39.3  # 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, one int32 op a lane a clock
HBM_PEAK_GBS = 8000.0
OPS_PER_CELL = 12  # a gap-affine cell's 7 integer ops for its score and 5 for its decision: the counting convention
CLASS_NAMES = ["sweep_group<1>", "sweep_group<2>", "sweep_group<4>", "sweep_group<8>", "sweep_wave<1>",
               "sweep_wave<2>", "sweep_wave<4>", "sweep_wave<8>", "sweep_wave<16>", "sweep_wave<32>"]
CLASS_WIDTHS = [32, 64, 128, 256, 64, 128, 256, 512, 1024, 2048]
CLASSES = len(CLASS_NAMES)


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_band_cells(n: np.ndarray, m: np.ndarray, cost: np.ndarray) -> np.ndarray:
    """Cells of the narrowest band the strict certificate proves for each pair at its true cost: lane diagonals
    (query less reference) low..high around the start and end diagonals, out to where a path off them must cost
    more (reference lanes.mojo:333-345,1370-1374), counted within the matrix."""
    end = m - n
    lo0, hi0 = np.minimum(0, end), np.maximum(0, end)

    def off(k, e):
        k = np.asarray(k)
        up = k > np.maximum(0, e)
        down = k < np.minimum(0, e)
        ins = lambda x: np.where(x > 0, OPEN + x * EXTEND, 0)
        dele = lambda x: np.where(x > 0, OPEN + x * EXTEND, 0)
        return np.where(up, ins(k) + dele(k - e), np.where(down, dele(-k) + ins(e - k), 0))

    high = hi0.copy()
    active = np.ones_like(high, dtype=bool)
    while active.any():
        grow = active & (high < m) & (off(high + 1, end) <= cost)
        high = np.where(grow, high + 1, high)
        active = grow
    low = lo0.copy()
    active = np.ones_like(low, dtype=bool)
    while active.any():
        grow = active & (low > -n) & (off(low - 1, end) <= cost)
        low = np.where(grow, low - 1, low)
        active = grow
    cells = np.zeros(len(n), dtype=np.float64)
    for r in range(1, int(n.max()) + 1):
        live = n >= r
        first = np.maximum(1, r + low)
        last = np.minimum(m, r + high)
        cells += np.where(live, np.maximum(0, last - first + 1), 0)
    return cells


def compile_native(name: str):
    """A restated CPU batch path (oracle/<name>.cpp), compiled for this host; the portable prebuilt one if that
    fails."""
    source = os.path.join(ROOT, "oracle", f"{name}.cpp")
    target = os.path.join(tempfile.gettempdir(), f"dinara_{name}_native_{os.getpid()}.so")
    path = target
    try:
        subprocess.run(["g++", "-O3", "-march=native", "-fPIC", "-std=c++17", "-pthread", "-shared", "-o", target,
                        source], check=True, capture_output=True, timeout=120)
        flags = "-O3 -march=native"
    except (OSError, subprocess.SubprocessError):
        path = os.path.join(ROOT, "oracle", f"lib{name}.so")
        flags = "-O3 -march=x86-64-v3 (prebuilt)"
    return ctypes.CDLL(path), flags


def load_baseline(cores: int):
    """The restated CPU batch path for costs."""
    lib, flags = compile_native("baseline")
    i64, vp = ctypes.c_int64, ctypes.c_void_p
    lib.baseline_alignments.argtypes = [i64, vp, vp, vp, vp, i64, i64, i64, i64, vp, vp, i64]
    lib.baseline_alignments.restype = i64
    return lib, flags


LOCAL_OPS_PER_CELL = 10  # a local cell: 4 adds, 5 maxes and the floor at zero, the counting convention


def run_local(args) -> int:
    """Local alignments under a substitution table: the local end, the reach back to its start and the span's
    global alignment with its walk, every pass on the device with the pairs resident in HBM."""
    t0 = time.time()
    pairs = args.pairs or 50_000
    alphabet, table = synthetic.protein_table()
    refs, qrys = synthetic.local_protein(pairs)
    n = np.fromiter((len(r) for r in refs), dtype=np.int64, count=pairs)
    m = np.fromiter((len(q) for q in qrys), dtype=np.int64, count=pairs)
    print(f"[bench] local workload ready: {pairs} pairs in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    handle = _lib.lib()
    if handle.dinara_device_count() < 1:
        print("[bench] no GPU visible", file=sys.stderr)
        return 1
    cores = host_cores()
    R, Q = _lib.Texts(refs), _lib.Texts(qrys)
    raw = (ctypes.c_int8 * len(table))(*table)
    scoring = _lib.CScoring(alphabet.encode(), len(alphabet), ctypes.cast(raw, ctypes.c_void_p), -11, -1)
    mode = _lib.CMode(2, 0, 0, 0, 0, 0, 0, 0, 0)
    status = ctypes.c_int64(0)
    t0 = time.time()
    batch = handle.dinara_scoring_batch_new(0, _lib.BATCH_ALIGNMENTS, pairs, *R.args(), *Q.args(),
                                            ctypes.byref(scoring), ctypes.byref(mode), 1, min(cores, 16),
                                            ctypes.byref(status))
    if not batch:
        print(f"[bench] dinara_scoring_batch_new failed: {status.value}", file=sys.stderr)
        return 1
    pack_s = time.time() - t0
    for _ in range(args.warmup):
        if handle.dinara_scoring_batch_run(batch) != 0:
            print("[bench] warm-up run failed", file=sys.stderr)
            return 1
    walls, passes = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        code = handle.dinara_scoring_batch_run(batch)
        walls.append(time.perf_counter() - t0)
        if code != 0:
            print(f"[bench] run failed: {code}", file=sys.stderr)
            return 1
        passes.append([handle.dinara_scoring_batch_stat(batch, w) for w in (0, 11, 12, 13)])
    print(f"[bench] {args.steps} steps, {np.mean(walls) * 1e3:.2f} ms each", file=sys.stderr, flush=True)
    found = (_lib.CAlignment * pairs)()
    statuses = (ctypes.c_int64 * pairs)()
    handle.dinara_scoring_batch_alignments(batch, ctypes.cast(R.pointers, ctypes.c_void_p),
                                           ctypes.cast(Q.pointers, ctypes.c_void_p), found, statuses)
    gpu = []
    for i in range(pairs):
        a = found[i]
        gpu.append((a.score, ctypes.string_at(a.cigar, a.cigar_length) if a.cigar else b"",
                    (a.reference_start, a.reference_end, a.query_start, a.query_end)) if statuses[i] == 0 else None)
        handle.dinara_free(a.cigar)
    handle.dinara_scoring_batch_free(batch)

    ms_per_step = float(np.mean(walls) * 1e3)
    mean = np.mean(np.array(passes), axis=0)
    cells = float((n * m).sum())
    # The local-end pass sweeps every cell of every pair in one launch (every query takes the same columns a lane).
    local_ms = float(mean[1])
    achieved = cells * LOCAL_OPS_PER_CELL / (local_ms / 1e3) / 1e12
    roofline = {"bound": "valu", "kernel": "score_wave<local>", "achieved": round(achieved, 3),
                "peak": VALU_PEAK_TOPS, "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
                "traffic": None, "launch_ms": round(local_ms, 4),
                "note": "full-matrix cells x 10 int ops (local Gotoh), over the local-end pass's one launch"}
    cpu, exact = None, None
    if not args.no_cpu:
        lib, flags = compile_native("baseline_scoring")
        i64, vp = ctypes.c_int64, ctypes.c_void_p
        lib.baseline_local_alignments.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp, i64, i64, i64, i64, vp, vp, vp,
                                                  i64]
        codes = {ord(ch): i for i, ch in enumerate(alphabet)}
        lut = np.zeros(256, dtype=np.uint8)
        for byte, code in codes.items():
            lut[byte] = code
        rc = [lut[np.frombuffer(r, dtype=np.uint8)].tobytes() for r in refs]
        qc = [lut[np.frombuffer(q, dtype=np.uint8)].tobytes() for q in qrys]
        RC, QC = _lib.Texts(rc), _lib.Texts(qc)
        size = len(alphabet)
        table32 = np.zeros(32 * 32, dtype=np.int8)
        for a in range(size):
            table32[a * 32:a * 32 + size] = table[a * size:(a + 1) * size]
        scores = (ctypes.c_int64 * pairs)()
        spans = (ctypes.c_int64 * (4 * pairs))()
        stride = int((n + m).max()) * 2 + 2
        cigars = ctypes.create_string_buffer(pairs * stride)
        t0 = time.perf_counter()
        lib.baseline_local_alignments(pairs, ctypes.cast(R.pointers, vp), ctypes.cast(R.lengths, vp),
                                      ctypes.cast(Q.pointers, vp), ctypes.cast(Q.lengths, vp),
                                      ctypes.cast(RC.pointers, vp), ctypes.cast(QC.pointers, vp),
                                      table32.ctypes.data, -12, -1, 1, cores, scores, spans, cigars, stride)
        cpu_s = time.perf_counter() - t0
        blob = cigars.raw
        same = 0
        for i in range(pairs):
            want = (scores[i], blob[i * stride:(i + 1) * stride].split(b"\0", 1)[0], tuple(spans[4 * i:4 * i + 4]))
            same += int(gpu[i] == want)
        exact = same / pairs
        cpu = {"value": round(pairs / cpu_s, 1), "unit": "pairs/s", "cores": cores, "kind": "port",
               "sample": f"the full workload ({pairs} pairs), one pass: local end in 16-bit lanes x 32 pairs, reach "
                         f"back and span traceback per pair, {flags}, {cpu_model()}"}
        print(f"[bench] CPU baseline {pairs / cpu_s:.0f} pairs/s on {cores} cores; GPU==CPU on {same}/{pairs}",
              file=sys.stderr, flush=True)
    line = {
        "metric": "pairs/s (local Scoring scores + CIGARs + spans, bit-exact)",
        "value": round(pairs / (ms_per_step / 1e3), 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"local alignment, 20-letter table (diag +4..+11, off -4..+3), gap -11/-1, {pairs} "
                               "pairs: 300 aa query at 30% mutation inside a 3000 aa window", "pairs": pairs},
        "gcups": round(cells / (ms_per_step / 1e3) / 1e9, 2),
        "device_ms_per_step": round(float(mean[0]), 3),
        "passes_ms_per_step": {"local_end": round(float(mean[1]), 4), "reach_back": round(float(mean[2]), 4),
                               "span_trace_and_walk": round(float(mean[3]), 4)},
        "bit_exact_vs_cpu": exact,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pack_upload_s": round(pack_s, 3),
    }
    print(json.dumps(line))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=None, help="pairs (default: the workload's own count)")
    ap.add_argument("--workload", choices=("affine", "local"), default="affine",
                    help="affine: global gap-affine (4,6,2) CIGARs, 100k x 250 bp (the default); local: local "
                         "alignment under a 20-letter table, 50k x (3000 aa window, 300 aa query)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (pairs would shard by length-balanced "
              "ranges, one process per GPU); run with --gpus 1", file=sys.stderr)
        return 2
    if args.workload == "local":
        return run_local(args)
    args.pairs = args.pairs or 100_000

    t0 = time.time()
    refs, qrys = synthetic.affine_reads(args.pairs)
    pairs = len(refs)
    n = np.fromiter((len(r) for r in refs), dtype=np.int64, count=pairs)
    m = np.fromiter((len(q) for q in qrys), dtype=np.int64, count=pairs)
    print(f"[bench] workload ready: {pairs} pairs in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    handle = _lib.lib()
    if handle.dinara_device_count() < 1:
        print("[bench] no GPU visible", file=sys.stderr)
        return 1
    cores = host_cores()
    R, Q = _lib.Texts(refs), _lib.Texts(qrys)
    costs = _lib.CCosts(MISMATCH, OPEN, EXTEND, -1, 0, 0, 0, -1, 0)
    options = _lib.COptions(-(1 << 63), (1 << 63) - 1, -1, 1, 0, 0)
    status = ctypes.c_int64(0)
    t0 = time.time()
    batch = handle.dinara_batch_new(0, _lib.BATCH_ALIGNMENTS, pairs, *R.args(), *Q.args(), ctypes.byref(costs),
                                    None, ctypes.byref(options), min(cores, 16), ctypes.byref(status))
    if not batch:
        print(f"[bench] dinara_batch_new failed: {status.value}", file=sys.stderr)
        return 1
    pack_s = time.time() - t0
    print(f"[bench] packed and uploaded in {pack_s:.2f}s", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        if handle.dinara_batch_run(batch) != 0:
            print("[bench] warm-up run failed", file=sys.stderr)
            return 1
    walls, device_ms, walk_ms, class_ms, class_launches, class_pairs = [], [], [], None, None, None
    for step in range(args.steps):
        t0 = time.perf_counter()
        code = handle.dinara_batch_run(batch)
        walls.append(time.perf_counter() - t0)
        if code != 0:
            print(f"[bench] run failed: {code}", file=sys.stderr)
            return 1
        device_ms.append(handle.dinara_batch_stat(batch, 0))
        walk_ms.append(handle.dinara_batch_stat(batch, 5))
        ms = np.array([handle.dinara_batch_stat(batch, 100 + c) for c in range(CLASSES)])
        class_ms = ms if class_ms is None else class_ms + ms
        class_launches = np.array([handle.dinara_batch_stat(batch, 200 + c) for c in range(CLASSES)])
        class_pairs = np.array([handle.dinara_batch_stat(batch, 400 + c) for c in range(CLASSES)])
    first_proven = handle.dinara_batch_stat(batch, 4)
    print(f"[bench] {args.steps} steps, {np.mean(walls) * 1e3:.2f} ms each", file=sys.stderr, flush=True)

    found = (_lib.CAlignment * pairs)()
    statuses = (ctypes.c_int64 * pairs)()
    handle.dinara_batch_alignments(batch, found, statuses)
    gpu_costs = np.array([found[i].cost if statuses[i] == 0 else -1 for i in range(pairs)], dtype=np.int64)
    gpu_cigars = [ctypes.string_at(found[i].cigar, found[i].cigar_length) if statuses[i] == 0 else b""
                  for i in range(pairs)]
    for i in range(pairs):
        handle.dinara_free(found[i].cigar)
    handle.dinara_batch_free(batch)

    ms_per_step = float(np.mean(walls) * 1e3)
    value = pairs / (ms_per_step / 1e3)
    gcups = float((n * m).sum()) / (ms_per_step / 1e3) / 1e9

    # Roofline of the dominant kernel: its algorithmic work (each pair's proven-band cells at the true cost, 12
    # integer ops a cell) over its launches' average time from the HIP events on the library's stream.
    mean_class_ms = class_ms / args.steps
    dominant = int(np.argmax(mean_class_ms))
    band_cells = oracle_band_cells(n, m, np.where(gpu_costs >= 0, gpu_costs, 0))
    launches = max(class_launches[dominant], 1)
    share = class_pairs[dominant] / pairs  # the dominant class sweeps this share of the pairs
    launch_ms = mean_class_ms[dominant] / launches
    algorithmic_ops = float(band_cells.sum()) * share * OPS_PER_CELL / launches
    achieved = algorithmic_ops / (launch_ms / 1e3) / 1e12
    width = CLASS_WIDTHS[dominant]
    swept_cells = float(np.mean(n)) * width * class_pairs[dominant] / launches
    swept_ops = swept_cells * OPS_PER_CELL / (launch_ms / 1e3) / 1e12
    roofline = {
        "bound": "valu",
        "kernel": CLASS_NAMES[dominant],
        "achieved": round(achieved, 3),
        "peak": VALU_PEAK_TOPS,
        "unit": "TOP/s",
        "frac": round(achieved / VALU_PEAK_TOPS, 4),
        "traffic": None,
        "launch_ms": round(launch_ms, 4),
        "swept_cells_rate_tops": round(swept_ops, 3),
        "note": "int32 VALU peak (no int MFMA in this path); achieved counts proven-band cells x 12 ops",
    }

    cpu = None
    exact = None
    if not args.no_cpu:
        lib, flags = load_baseline(cores)
        stride = int((n + m).max()) * 2 + 2
        cpu_costs = (ctypes.c_int64 * pairs)()
        cigars = ctypes.create_string_buffer(pairs * stride)
        t0 = time.perf_counter()
        lib.baseline_alignments(pairs, ctypes.cast(R.pointers, ctypes.c_void_p),
                                ctypes.cast(R.lengths, ctypes.c_void_p), ctypes.cast(Q.pointers, ctypes.c_void_p),
                                ctypes.cast(Q.lengths, ctypes.c_void_p), MISMATCH, OPEN, EXTEND, cores, cpu_costs,
                                cigars, stride)
        cpu_s = time.perf_counter() - t0
        raw = cigars.raw
        same = 0
        for i in range(pairs):
            cigar = raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0]
            same += int(cpu_costs[i] == gpu_costs[i] and cigar == gpu_cigars[i])
        exact = same / pairs
        cpu = {
            "value": round(pairs / cpu_s, 1),
            "unit": "pairs/s",
            "cores": cores,
            "kind": "port",
            "sample": f"the full workload ({pairs} pairs), one pass, 16-bit lanes x 32 pairs, {flags}, "
                      f"{cpu_model()}",
        }
        print(f"[bench] CPU baseline {pairs / cpu_s:.0f} pairs/s on {cores} cores; GPU==CPU on {same}/{pairs}",
              file=sys.stderr, flush=True)

    line = {
        "metric": "pairs/s (global gap-affine scores + Ties.LEFT CIGARs, bit-exact)",
        "value": round(value, 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": "global gap-affine (4,6,2) scores + CIGARs, 100,000 pairs of 250 bp at 5% "
                        "(60/20/20, indels Geometric(0.7)), 2-bit packed",
            "pairs": pairs,
            "initial_band": int(os.environ.get("DINARA_INITIAL_WIDTH", "32")),
        },
        "gcups": round(gcups, 2),
        "device_ms_per_step": round(float(np.mean(device_ms)), 3),
        "first_band_proved": round(first_proven / pairs, 4),
        "bit_exact_vs_cpu": exact,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels_ms_per_step": {**{CLASS_NAMES[c]: round(float(mean_class_ms[c]), 4) for c in range(CLASSES)
                                   if mean_class_ms[c] > 0}, "walk_bands": round(float(np.mean(walk_ms)), 4)},
        "pairs_per_class": {CLASS_NAMES[c]: int(class_pairs[c]) for c in range(CLASSES) if class_pairs[c] > 0},
        "pack_upload_s": round(pack_s, 3),
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())

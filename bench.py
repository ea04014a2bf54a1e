#!/usr/bin/env python3
"""Benchmark: batched device alignment on one MI355X, learned band against plain band doubling.

Default workload (`--workload uneven`, one step): unit-cost distances and Ties.LEFT CIGARs of 200,000 pairs whose
lengths are log-normal around 1.5 kbp (sigma 1, clipped to 100 bp..20 kbp) at a per-pair error rate uniform in
1%..15% (half substitutions, a quarter each insertions and deletions). The pairs are packed and uploaded once
(dinara_batch_new); a step is the whole device path with its inputs resident in HBM. Three ways over the same
pairs, compared byte for byte:
  plain doubling   every band from 64 diagonals, twice as wide until the certificate proves it, each band keeping
                   its decisions; one pair a lane group, one launch after another;
  persistent       length-binned persistent launches replayed as one hipGraph (bins side by side on streams), first
                   bands and launch shape autotuned on the batch, then each proven band traced again;
  learned          the same launches with each pair's first band from a network trained (on the GPU) on the
                   aligner's own distances of separate pairs, seeded on the device inside the step, sweeps keeping
                   their decisions so a proven pair is walked at once; launch form, margin, slope and the first
                   pass's bin edges (which bins first bands may take) chosen on held-out pairs.
`value` is the learned path's pairs/s. The restated CPU path (oracle/baseline_edit.cpp, per-pair bit-parallel band
doubling) is timed on a sample of the pairs on all host cores and checked against the GPU.

Other workloads: `affine` (gap-affine (4,6,2) CIGARs, 100,000 x 250 bp), `edit` (unit-cost distances, 10,000 x
150 bp), `local` (local alignment under a 20-letter table), `learned` (gap-affine distances of uneven pairs, the
learned band against plain doubling on the gap-cost sweeps).

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

# HIP maps a process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default): the bins that run side by side
# on their own streams need more to run at once. Read when the library first touches the GPU; DINARA_HW_QUEUES
# chooses another count (4 for HIP's default).
os.environ["GPU_MAX_HW_QUEUES"] = os.environ.get("DINARA_HW_QUEUES", "16")

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
               "sweep_wave<2>", "sweep_wave<4>", "sweep_wave<8>", "sweep_wave<16>", "sweep_wave<32>",
               "sweep_block<4>x1", "sweep_block<8>x1", "sweep_block<4>x3", "sweep_block<8>x2", "sweep_block<8>x3",
               "sweep_block<8>x4", "sweep_block<8>x5", "sweep_block<8>x6", "sweep_block<8>x7", "sweep_block<8>x8",
               "sweep_block<12>x8", "sweep_block<16>x8"]
CLASS_WIDTHS = [32, 64, 128, 256, 64, 128, 256, 512, 1024, 2048, 512, 1024, 1536, 2048, 3072, 4096, 5120, 6144, 7168,
                8192, 12288, 16384]
CLASSES = len(CLASS_NAMES)
# HBM bytes per launch of the dominant kernel, from rocprofv3 FETCH_SIZE and WRITE_SIZE passes, each scaled by what
# the same counter reads for a plain 1 GiB device copy: FETCH_SIZE x 2.0, WRITE_SIZE x 1.0. The affine step (100,000
# pairs of 250 bp): profiles/sweep_group_affine_100k_x_250_hbm_counters.md. The learned run of the mixed-length
# workload (200,000 pairs), every edit_sweep dispatch of one run summed (the bins run side by side):
# profiles/edit_sweep_learned_uneven_200k_hbm_counters.md.
MEASURED_TRAFFIC = {("affine", 100_000, "sweep_group<2>"): {"read": 119.1e6, "written": 823.3e6},
                    ("uneven", 200_000, "edit_sweep"): {"read": 2.439e9, "written": 87.46e9}}


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


def _sum_min_const_minus(P, Q, a, b):
    """sum over k = a..b of min(P, Q - k), elementwise (0 for an empty range)."""
    s = Q - P
    n1 = np.maximum(0, np.minimum(b, s) - a + 1)
    lo2 = np.maximum(a, s + 1)
    n2 = np.maximum(0, b - lo2 + 1)
    return np.where(b >= a, n1 * P + n2 * Q - np.where(n2 > 0, (lo2 + b) * n2 // 2, 0), 0)


def _sum_min_plus_const(P, R, a, b):
    """sum over k = a..b of min(P + k, R), elementwise."""
    s = R - P
    hi1 = np.minimum(b, s - 1)
    n1 = np.maximum(0, hi1 - a + 1)
    lo2 = np.maximum(a, s)
    n2 = np.maximum(0, b - lo2 + 1)
    return np.where(b >= a, n1 * P + np.where(n1 > 0, (a + hi1) * n1 // 2, 0) + n2 * R, 0)


def band_cells(n, m, low, high):
    """Cells (r, c) of an n x m matrix, 1 <= r <= n, 1 <= c <= m, on diagonals low <= c - r <= high (closed form:
    a diagonal k >= 0 holds min(n, m - k) cells, k < 0 min(n + k, m))."""
    n = n.astype(np.int64)
    m = m.astype(np.int64)
    lo = np.maximum(low, -(n - 1))
    hi = np.minimum(high, m - 1)
    a = np.maximum(lo, 0)
    pos = np.where(hi >= a, _sum_min_const_minus(n, m, a, hi), 0)
    b2 = np.minimum(hi, -1)
    neg = np.where(b2 >= lo, _sum_min_plus_const(n, m, lo, b2), 0)
    return pos + neg


def proven_band_cells(n, m, cost, strict, opening=OPEN, extension=EXTEND):
    """The cells of the narrowest band the certificate proves at each pair's true cost under gap costs (opening,
    extension), (4, 6, 2) unless given: out from the start and end diagonals while a path off the band could cost
    less (no more, `strict`) than the pair."""
    n = n.astype(np.int64)
    m = m.astype(np.int64)
    cost = cost.astype(np.int64)
    end = m - n

    def off(k):
        up = k > np.maximum(0, end)
        down = k < np.minimum(0, end)
        gap = lambda x: np.where(x > 0, opening + x * extension, 0)
        return np.where(up, gap(k) + gap(k - end), np.where(down, gap(-k) + gap(end - k), 0))

    def grow(start, step, limit):
        # The furthest k from `start` (by `step`) whose off-band price still admits the pair, binary search.
        span = np.zeros_like(start)
        bit = 1 << 16
        while bit:
            trial = span + bit
            k = start + step * trial
            inside = (trial <= limit) & ((off(k) <= cost) if strict else (off(k) < cost))
            span = np.where(inside, trial, span)
            bit >>= 1
        return start + step * span

    high = grow(np.maximum(0, end), 1, m + 1)
    low = grow(np.minimum(0, end), -1, n + 1)
    return band_cells(n, m, low, high)


def run_learned(args) -> int:
    """Gap-affine (4, 6, 2) distances of uneven pairs: the plain path (each pair's first band its start and end
    diagonals, then the band its found cost proves) against the learned path (the network trained on the aligner's
    own costs seeds each pair's first band), on the same pairs in the same run, compared byte for byte."""
    train_n = args.train or 1_000_000
    eval_n = args.pairs or 200_000
    cores = host_cores()
    workers = min(cores, 16)
    held_n = max(eval_n // 2, 1)
    t0 = time.time()
    train_r, train_q = synthetic.uneven_parallel(train_n, seed=41, workers=workers)
    eval_r, eval_q = synthetic.uneven_parallel(eval_n, seed=42, workers=workers)
    held_r, held_q = synthetic.uneven_parallel(held_n, seed=43, workers=workers)
    print(f"[bench] uneven workload ready: {train_n} training, {held_n} held-out, {eval_n} evaluation pairs in "
          f"{time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    handle = _lib.lib()
    if handle.dinara_device_count() < 1:
        print("[bench] no GPU visible", file=sys.stderr)
        return 1
    costs = _lib.CCosts(MISMATCH, OPEN, EXTEND, -1, 0, 0, 0, -1, 0)
    F = 288

    def batch_of(refs, qrys):
        R, Q = _lib.Texts(refs), _lib.Texts(qrys)
        status = ctypes.c_int64(0)
        b = handle.dinara_batch_new(0, _lib.BATCH_DISTANCES, len(refs), *R.args(), *Q.args(), ctypes.byref(costs),
                                    None, None, workers, ctypes.byref(status))
        if not b:
            raise RuntimeError(f"dinara_batch_new failed: {status.value}")
        return b, (R, Q)

    def distances(b, count):
        out = (ctypes.c_int64 * count)()
        if handle.dinara_batch_distances(b, out) != 0:
            raise RuntimeError("distances failed")
        return np.array(out[:], dtype=np.int64)

    def features(b, count):
        x = np.zeros((count, F), dtype=np.float32)
        if handle.dinara_batch_features(b, x.ctypes.data) != 0:
            raise RuntimeError("features failed")
        return x

    # The exact aligner labels its own training pairs.
    t0 = time.time()
    tb, keep = batch_of(train_r, train_q)
    if handle.dinara_batch_run(tb) != 0:
        raise RuntimeError("labelling run failed")
    labels = distances(tb, train_n)
    x_train = features(tb, train_n)
    handle.dinara_batch_free(tb)
    del keep, train_r, train_q
    y_train = np.log1p(np.maximum(labels, 0)).astype(np.float32)
    label_s = time.time() - t0
    net = handle.dinara_mlp_new(0, 17)
    loss = ctypes.c_double(0)
    t0 = time.time()
    epochs = args.epochs or 3
    if handle.dinara_mlp_train(net, train_n, x_train.ctypes.data, y_train.ctypes.data, epochs, 4096, 2e-3, 7,
                               ctypes.byref(loss)) != 0:
        raise RuntimeError("training failed")
    train_s = time.time() - t0
    print(f"[bench] labelled in {label_s:.1f}s, trained {epochs} epochs in {train_s:.1f}s, loss {loss.value:.4f}",
          file=sys.stderr, flush=True)
    del x_train

    def timed(b):
        for _ in range(args.warmup):
            if handle.dinara_batch_run(b) != 0:
                raise RuntimeError("warm-up failed")
        walls, dev, cls = [], [], np.zeros(CLASSES)
        for _ in range(args.steps):
            t = time.perf_counter()
            if handle.dinara_batch_run(b) != 0:
                raise RuntimeError("run failed")
            walls.append(time.perf_counter() - t)
            dev.append(handle.dinara_batch_stat(b, 0))
            cls += np.array([handle.dinara_batch_stat(b, 100 + c) for c in range(CLASSES)])
        launches = np.array([handle.dinara_batch_stat(b, 200 + c) for c in range(CLASSES)])
        first = handle.dinara_batch_stat(b, 4)
        return float(np.mean(walls) * 1e3), float(np.mean(dev)), cls / args.steps, launches, first

    def path(doubling, margin, refs=eval_r, qrys=eval_q):
        b, keep = batch_of(refs, qrys)
        handle.dinara_batch_set_retry(b, 1 if doubling else 0)
        seed_ms = 0.0
        if margin is not None:
            t = time.perf_counter()
            if handle.dinara_batch_seed_learned(b, net, margin) != 0:
                raise RuntimeError("seeding failed")
            seed_ms = (time.perf_counter() - t) * 1e3
        ms, dev, cls, launches, first = timed(b)
        out = distances(b, len(refs))
        handle.dinara_batch_free(b)
        del keep
        return {"ms": ms, "dev": dev, "cls": cls, "launches": launches, "first": first, "seed_ms": seed_ms,
                "out": out}

    # The seeds' margin is chosen on held-out pairs drawn like the workload, then frozen for the evaluation pairs;
    # the certified path is timed on them too.
    held = {margin: path(False, margin, held_r, held_q) for margin in LEARNED_MARGINS}
    held_certified = path(False, None, held_r, held_q)
    del held_r, held_q
    best_margin = min(held, key=lambda k: held[k]["ms"] + held[k]["seed_ms"])
    print("[bench] held-out: " + ", ".join(f"margin {k}: {v['ms'] + v['seed_ms']:.1f} ms" for k, v in held.items()) +
          f", certified {held_certified['ms']:.1f} ms", file=sys.stderr, flush=True)
    doubling = path(True, None)
    certified = path(False, None)
    learned = path(False, best_margin)
    handle.dinara_mlp_free(net)
    exact = float(np.mean([np.array_equal(doubling["out"], r["out"]) for r in (certified, learned)]))
    print(f"[bench] doubling {doubling['ms']:.1f} ms, certified {certified['ms']:.1f} ms, learned "
          f"{learned['ms'] + learned['seed_ms']:.1f} ms (margin {best_margin})", file=sys.stderr, flush=True)

    # The CPU baseline: the restated lanes path (oracle/baseline.cpp, 16-bit lanes, distances) on a sample of the
    # evaluation pairs, those whose dearest path its 16-bit cells hold (lanes.mojo:317-331).
    cpu = None
    if not args.no_cpu:
        lib, flags = load_baseline(cores)
        rng = np.random.default_rng(5)
        held_cells = lambda i: 2 * OPEN + (len(eval_r[i]) + len(eval_q[i])) * EXTEND + max(MISMATCH, OPEN + EXTEND)
        pool = [int(i) for i in rng.permutation(eval_n) if held_cells(int(i)) < 16000]
        sample = pool[:min(len(pool), args.cpu_sample or 20_000)]
        Rs, Qs = _lib.Texts([eval_r[i] for i in sample]), _lib.Texts([eval_q[i] for i in sample])
        cpu_costs = (ctypes.c_int64 * len(sample))()
        t0 = time.perf_counter()
        lib.baseline_alignments(len(sample), ctypes.cast(Rs.pointers, ctypes.c_void_p),
                                ctypes.cast(Rs.lengths, ctypes.c_void_p), ctypes.cast(Qs.pointers, ctypes.c_void_p),
                                ctypes.cast(Qs.lengths, ctypes.c_void_p), MISMATCH, OPEN, EXTEND, cores, cpu_costs,
                                None, 0)
        cpu_s = time.perf_counter() - t0
        same = sum(int(cpu_costs[j] == learned["out"][i]) for j, i in enumerate(sample))
        cpu = {"value": round(len(sample) / cpu_s, 1), "unit": "pairs/s", "cores": cores, "kind": "port",
               "sample": f"{len(sample)} evaluation pairs drawn uniformly from the {len(pool)} whose dearest path "
                         f"the lanes path's 16-bit cells hold (of {eval_n}), distances, 16-bit lanes x 32 pairs, "
                         f"{flags}, {cpu_model()}", "equal_to_gpu": round(same / max(len(sample), 1), 4)}
        print(f"[bench] CPU baseline {len(sample) / cpu_s:.0f} pairs/s on {cores} cores; equal on "
              f"{same}/{len(sample)}", file=sys.stderr, flush=True)
    n = np.fromiter((len(r) for r in eval_r), dtype=np.int64, count=eval_n)
    m = np.fromiter((len(q) for q in eval_q), dtype=np.int64, count=eval_n)
    cells = proven_band_cells(n, m, np.maximum(learned["out"], 0), strict=False)
    learned_cls, learned_launches = learned["cls"], learned["launches"]
    dominant = int(np.argmax(learned_cls))
    launch_ms = learned_cls[dominant] / max(learned_launches[dominant], 1)
    achieved = float(cells.sum()) * 7 / (float(learned_cls.sum()) / 1e3) / 1e12
    roofline = {"bound": "valu", "kernel": CLASS_NAMES[dominant], "achieved": round(achieved, 3),
                "peak": VALU_PEAK_TOPS, "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4), "traffic": None,
                "launch_ms": round(float(launch_ms), 4),
                "note": "proven-band cells at the true cost x 7 int ops (distances keep no decisions), over every "
                        "sweep kernel's time of the learned run"}
    learned_total = learned["ms"] + learned["seed_ms"]
    line = {
        "metric": "pairs/s (gap-affine distances, uneven lengths): learned band vs plain band doubling, bit-exact",
        "value": round(eval_n / (learned_total / 1e3), 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(learned_total, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"gap-affine (4,6,2) distances, {eval_n} pairs log-normal around 1.5 kbp (sigma 1, "
                               f"100 bp..20 kbp), 1-15% error (50/25/25); network trained on {train_n} pairs "
                               "labelled by the aligner, its guesses the first bands", "pairs": eval_n,
                   "train_pairs": train_n, "epochs": epochs, "margin": best_margin},
        "plain_doubling": {"value": round(eval_n / (doubling["ms"] / 1e3), 1), "ms_per_step": round(doubling["ms"], 3),
                           "first_band_proved": round(doubling["first"] / eval_n, 4)},
        "plain_certified": {"value": round(eval_n / (certified["ms"] / 1e3), 1),
                            "ms_per_step": round(certified["ms"], 3),
                            "first_band_proved": round(certified["first"] / eval_n, 4)},
        "learned": {"ms_per_step": round(learned["ms"], 3), "seed_ms": round(learned["seed_ms"], 3),
                    "first_band_proved": round(learned["first"] / eval_n, 4),
                    "margin_chosen_on": f"{held_n} held-out pairs drawn like the workload (seed 43)",
                    "by_margin_ms_held_out": {str(k): round(v["ms"] + v["seed_ms"], 3) for k, v in held.items()},
                    "certified_ms_held_out": round(held_certified["ms"], 3)},
        "speedup_vs_plain_doubling": round(doubling["ms"] / learned_total, 3),
        "speedup_vs_plain_certified": round(certified["ms"] / learned_total, 3),
        "gcups": round(float((n * m).sum()) / (learned_total / 1e3) / 1e9, 2),
        "bit_exact_vs_plain": exact,
        "training": {"label_s": round(label_s, 2), "train_s": round(train_s, 2), "final_loss": round(loss.value, 5)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels_ms_per_step": {name: {CLASS_NAMES[c]: round(float(r["cls"][c]), 3) for c in range(CLASSES)
                                       if r["cls"][c] > 0}
                                for name, r in (("doubling", doubling), ("certified", certified),
                                                ("learned", learned))},
    }
    print(json.dumps(line))
    return 0


# Seeds e^(prediction + margin) - 1: room, or not, for the network's error. Below zero, a first band that fails
# still finds a cost near the pair's, whose certified band is then about the narrowest that proves it.
LEARNED_MARGINS = (-0.6, -0.4, -0.2, 0.0, 0.2)

FULL_OPS_PER_CELL = 14  # a whole-matrix cell: 3 adds and 5 mins of the recurrence, 6 ops of its decision nibble


def run_oversized(args) -> int:
    """Gap-affine (4, 6, 2) CIGARs of pairs of 70 kbp at about 1% edits: 140,000 letters a pair, past what every band
    kernel stages in LDS, so each takes the whole-matrix path (full.hip) after the band sweeps leave it. The CPU
    baseline is the oracle's proven-band trace (oracle.c, one core) of one of the pairs."""
    count = args.pairs or 4
    rng = np.random.default_rng(70)
    letters = np.frombuffer(b"ACGT", dtype=np.uint8)
    refs, qrys = [], []
    for _ in range(count):
        r = letters[rng.integers(0, 4, 70_000)]
        q = r.copy()
        at = rng.choice(70_000, 700, replace=False)
        q[at] = letters[rng.integers(0, 4, 700)]  # about 0.75% substitutions
        cut = np.sort(rng.choice(69_000, 20, replace=False))
        pieces, last = [], 0
        for i, c in enumerate(cut):  # ten deletions and ten insertions of 1-8 letters
            pieces.append(q[last:c])
            if i % 2:
                pieces.append(letters[rng.integers(0, 4, int(rng.integers(1, 9)))])
                last = c
            else:
                last = c + int(rng.integers(1, 9))
        pieces.append(q[last:])
        refs.append(r.tobytes())
        qrys.append(np.concatenate(pieces).tobytes())
    handle = _lib.lib()
    R, Q = _lib.Texts(refs), _lib.Texts(qrys)
    costs = _lib.CCosts(MISMATCH, OPEN, EXTEND, -1, 0, 0, 0, -1, 0)
    status = ctypes.c_int64(0)
    b = handle.dinara_batch_new(0, _lib.BATCH_ALIGNMENTS, count, *R.args(), *Q.args(), ctypes.byref(costs), None,
                                None, 16, ctypes.byref(status))
    if not b:
        raise RuntimeError(f"dinara_batch_new failed: {status.value}")
    for _ in range(args.warmup):
        if handle.dinara_batch_run(b) != 0:
            raise RuntimeError("warm-up failed")
    walls, whole = [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        if handle.dinara_batch_run(b) != 0:
            raise RuntimeError("run failed")
        walls.append((time.perf_counter() - t) * 1e3)
        whole.append(handle.dinara_batch_stat(b, 11))
    swept = handle.dinara_batch_stat(b, 10)
    found = (_lib.CAlignment * count)()
    statuses = (ctypes.c_int64 * count)()
    if handle.dinara_batch_alignments(b, found, statuses) != 0:
        raise RuntimeError("alignments failed")
    got = []
    for i in range(count):
        got.append((found[i].cost, ctypes.string_at(found[i].cigar, found[i].cigar_length)) if statuses[i] == 0
                   else (int(statuses[i]), b""))
        if statuses[i] == 0:
            handle.dinara_free(found[i].cigar)
    handle.dinara_batch_free(b)
    ms = float(np.mean(walls))
    whole_ms = float(np.mean(whole))
    cells = float(sum(len(r) * len(q) for r, q in zip(refs, qrys)))
    achieved = cells * FULL_OPS_PER_CELL / (whole_ms / 1e3) / 1e12
    print(f"[bench] {count} pairs of 70 kbp: {ms:.1f} ms a run, whole-matrix pass {whole_ms:.1f} ms over {swept:.0f} "
          f"pairs", file=sys.stderr, flush=True)
    cpu = None
    exact = None
    if not args.no_cpu:
        from oracle import oracle  # the checker, on the CPU leg only

        oc = oracle.Costs.affine(MISMATCH, OPEN, EXTEND)
        t0 = time.perf_counter()
        want = oracle.costs_align_proven(refs[0], qrys[0], oc)
        cpu_s = time.perf_counter() - t0
        exact = bool(want[0] == got[0][0] and want[1].encode() == got[0][1])
        cpu = {"value": round(1 / cpu_s, 4), "unit": "pairs/s", "cores": 1, "kind": "port",
               "sample": "one of the pairs: the oracle's trace of its strict proven band (oracle.c "
                         f"orc_costs_align_proven), {cpu_s:.1f} s", "equal_to_gpu": exact}
        print(f"[bench] CPU (oracle, proven band) {cpu_s:.1f} s for one pair; equal: {exact}", file=sys.stderr,
              flush=True)
    line = {
        "metric": "pairs/s (gap-affine CIGARs of 70 kbp pairs past every band kernel: the whole-matrix path)",
        "value": round(count / (ms / 1e3), 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{count} pairs of 70,000 x about 70,000 letters, about 1% edits, gap-affine (4,6,2) "
                               "CIGARs, Ties.LEFT", "pairs": count},
        "whole_matrix_ms": round(whole_ms, 3),
        "pairs_on_the_whole_matrix_path": swept,
        "gcups": round(cells / (whole_ms / 1e3) / 1e9, 2),
        "roofline": {"bound": "valu", "kernel": "full_sweep<1> + full_walk", "achieved": round(achieved, 3),
                     "peak": VALU_PEAK_TOPS, "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
                     "traffic": None, "launch_ms": round(whole_ms / count, 3),
                     "note": f"every cell of the matrix x {FULL_OPS_PER_CELL} int32 ops, over the whole-matrix pass "
                             "(sweep, walk and CIGAR text of each pair, host clock)"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


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


def edit_bins(handle, b):
    """The unit path's bins (the library's count) and their widths in diagonals, read from a batch."""
    count = int(handle.dinara_batch_stat(b, 12))
    return count, [int(handle.dinara_batch_stat(b, 1200 + i)) for i in range(count)]


BIT_OPS_FIRST = 34   # int32 ops a 64-cell word of the bit-parallel step (17 on 64-bit words)
BIT_OPS_TRACE = 56   # the same with the 2-bit decisions of the tracing sweep


def load_edit_baseline():
    lib, flags = compile_native("baseline_edit")
    i64, vp = ctypes.c_int64, ctypes.c_void_p
    lib.baseline_edit.argtypes = [i64, vp, vp, vp, vp, i64, i64, vp, vp, i64]
    lib.baseline_edit.restype = i64
    return lib, flags


def edit_batch(handle, refs, qrys, kind, threads):
    R, Q = _lib.Texts(refs), _lib.Texts(qrys)
    costs = _lib.CCosts(1, 0, 1, -1, 0, 0, 0, -1, 0)
    status = ctypes.c_int64(0)
    b = handle.dinara_batch_new(0, kind, len(refs), *R.args(), *Q.args(), ctypes.byref(costs), None, None, threads,
                                ctypes.byref(status))
    if not b:
        raise RuntimeError(f"dinara_batch_new failed: {status.value}")
    if handle.dinara_batch_stat(b, 6) != 1:
        raise RuntimeError("the unit-cost batch did not take the bit-parallel path")
    return b, (R, Q)


def edit_results(handle, b, kind, count):
    """(costs, CIGARs) of the last run, as arrays / bytes."""
    if kind == _lib.BATCH_DISTANCES:
        out = (ctypes.c_int64 * count)()
        if handle.dinara_batch_distances(b, out) != 0:
            raise RuntimeError("distances failed")
        return np.array(out[:], dtype=np.int64), None
    found = (_lib.CAlignment * count)()
    statuses = (ctypes.c_int64 * count)()
    if handle.dinara_batch_alignments(b, found, statuses) != 0:
        raise RuntimeError("alignments failed")
    costs = np.array([found[i].cost if statuses[i] == 0 else -1 for i in range(count)], dtype=np.int64)
    cigars = [ctypes.string_at(found[i].cigar, found[i].cigar_length) if statuses[i] == 0 else b""
              for i in range(count)]
    for i in range(count):
        if statuses[i] == 0:
            handle.dinara_free(found[i].cigar)
    return costs, cigars


def timed_steps(handle, b, steps, warmup, before=None):
    """Wall ms of each of `steps` runs after `warmup` (each run preceded by `before()` when given, inside the
    timed region), and the device ms the library's events saw."""
    for _ in range(warmup):
        if before:
            before()
        if handle.dinara_batch_run(b) != 0:
            raise RuntimeError("warm-up run failed")
    walls, device = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        if before:
            before()
        if handle.dinara_batch_run(b) != 0:
            raise RuntimeError("run failed")
        walls.append((time.perf_counter() - t0) * 1e3)
        device.append(handle.dinara_batch_stat(b, 0))
    return float(np.mean(walls)), float(np.mean(device))


def end_to_end_learned(handle, refs, qrys, workers, net, margin, slope, chosen, floor, use_graph, first_bins):
    """The learned path end to end on a fresh batch, as a caller holding the sequences in host memory meets it: pack
    and upload (dinara_batch_new), learned seeding and the run (the first run of a batch: its graph captured, its
    arena reserved), and the alignments copied out with their CIGARs malloc'ed (dinara_batch_alignments). Then the
    drop-in call alone (dinara_gpu_alignments: the default persistent path, no seeds) on the same pairs."""
    count = len(refs)
    R, Q = _lib.Texts(refs), _lib.Texts(qrys)
    costs = _lib.CCosts(1, 0, 1, -1, 0, 0, 0, -1, 0)
    status = ctypes.c_int64(0)
    found = (_lib.CAlignment * count)()
    statuses = (ctypes.c_int64 * count)()

    def release():
        for i in range(count):
            if statuses[i] == 0:
                handle.dinara_free(found[i].cigar)

    t0 = time.perf_counter()
    b = handle.dinara_batch_new(0, _lib.BATCH_ALIGNMENTS, count, *R.args(), *Q.args(), ctypes.byref(costs), None,
                                None, workers, ctypes.byref(status))
    if not b:
        raise RuntimeError(f"dinara_batch_new failed: {status.value}")
    t1 = time.perf_counter()
    handle.dinara_batch_set_tuning(b, chosen[0], (ctypes.c_int64 * 5)(0, 0, 0, *floor))
    handle.dinara_batch_set_bins(b, first_bins)
    handle.dinara_batch_set_side_by_side(b, chosen[6])
    handle.dinara_batch_set_keep_first(b, 1)
    handle.dinara_batch_set_graph(b, use_graph)
    if handle.dinara_batch_seed_learned_by_length(b, net, ctypes.c_double(margin), ctypes.c_double(slope)) != 0 or \
            handle.dinara_batch_run(b) != 0:
        raise RuntimeError("end-to-end run failed")
    t2 = time.perf_counter()
    if handle.dinara_batch_alignments(b, found, statuses) != 0:
        raise RuntimeError("end-to-end fetch failed")
    t3 = time.perf_counter()
    release()
    handle.dinara_batch_free(b)
    t4 = time.perf_counter()
    if handle.dinara_gpu_alignments(0, count, *R.args(), *Q.args(), ctypes.byref(costs), None, None, workers, found,
                                    statuses) != 0:
        raise RuntimeError("drop-in call failed")
    t5 = time.perf_counter()
    release()
    total = t3 - t0
    return {"pairs_per_s": round(count / total, 1), "ms": round(total * 1e3, 1),
            "pack_upload_ms": round((t1 - t0) * 1e3, 1), "seed_and_run_ms": round((t2 - t1) * 1e3, 1),
            "fetch_ms": round((t3 - t2) * 1e3, 1),
            "drop_in_call_pairs_per_s": round(count / (t5 - t4), 1), "drop_in_call_ms": round((t5 - t4) * 1e3, 1),
            "note": "learned path on a fresh batch: pack + upload, seed + first run (graph capture, arena), copy out "
                    "with CIGARs malloc'ed; the drop-in dinara_gpu_alignments call (default path, no seeds) beside it"}


def drop_in_call(handle, R, Q, count, costs, options, threads, reps=3, kind=_lib.BATCH_ALIGNMENTS):
    """The end-to-end rate of the drop-in call (dinara_gpu_alignments, or dinara_gpu_distances): validation, packing,
    upload, the run, the copy back and (alignments) each CIGAR malloc'ed, for pairs already in host memory; the
    mean of `reps` calls after an untimed one. Freeing the CIGARs is the caller's and not timed."""
    found = (_lib.CAlignment * count)()
    statuses = (ctypes.c_int64 * count)()
    results = (ctypes.c_int64 * count)()
    walls = []
    for rep in range(reps + 1):
        t0 = time.perf_counter()
        if kind == _lib.BATCH_ALIGNMENTS:
            code = handle.dinara_gpu_alignments(0, count, *R.args(), *Q.args(), costs, None, options, threads, found,
                                                statuses)
        else:
            code = handle.dinara_gpu_distances(0, count, *R.args(), *Q.args(), costs, None, options, threads, results)
        wall = time.perf_counter() - t0
        if code != 0:
            raise RuntimeError(f"drop-in call failed: {code}")
        if kind == _lib.BATCH_ALIGNMENTS:
            for i in range(count):
                if statuses[i] == 0:
                    handle.dinara_free(found[i].cigar)
        if rep:
            walls.append(wall)
    ms = float(np.mean(walls) * 1e3)
    return {"pairs_per_s": round(count / (ms / 1e3), 1), "ms": round(ms, 2), "threads": threads,
            "note": "the whole drop-in call on host-resident pairs: pack, upload, run, copy back"
                    + (", CIGARs malloc'ed" if kind == _lib.BATCH_ALIGNMENTS else "")}


def bin_profile(handle, b):
    """One run launched kernel by kernel (not the graph): each bin's first-pass and tracing time and sweeps."""
    handle.dinara_batch_set_graph(b, 0)
    if handle.dinara_batch_run(b) != 0:
        raise RuntimeError("profiling run failed")
    bins, widths = edit_bins(handle, b)
    prof = {k: [handle.dinara_batch_stat(b, base + i) for i in range(bins)]
            for k, base in (("sweep_ms", 600), ("sweeps", 700), ("trace_ms", 800), ("traces", 900),
                            ("sweep_word_columns", 1000), ("trace_word_columns", 1100))}
    prof["widths"] = widths
    prof["walk_ms"] = handle.dinara_batch_stat(b, 5)
    prof["run_ms"] = handle.dinara_batch_stat(b, 0)
    handle.dinara_batch_set_graph(b, 1)
    return prof


def run_uneven(args) -> int:
    """Unit-cost distances and Ties.LEFT CIGARs of an uneven batch (the mixed-length workload): plain GPU band
    doubling (each pair's band from 64 diagonals, twice as wide until proven; one pair a lane group, launches one
    by one) against the persistent length-binned path (lane groups refilling from each bin's queue, the run one
    captured hipGraph, start bins and workgroups autotuned on the batch) and against the same with learned band
    seeds (the network trained on the aligner's own distances of separate pairs). Every variant's results are
    compared byte for byte."""
    pairs = args.pairs or 200_000
    train_n = args.train if args.train is not None else 1_000_000
    cores = host_cores()
    workers = min(cores, 16)
    t0 = time.time()
    refs, qrys = synthetic.uneven_parallel(pairs, seed=3, workers=workers)
    n = np.fromiter((len(r) for r in refs), dtype=np.int64, count=pairs)
    m = np.fromiter((len(q) for q in qrys), dtype=np.int64, count=pairs)
    print(f"[bench] uneven workload ready: {pairs} pairs in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    # The training pairs and the held-out tuning pairs are made here too, before the first GPU call (no process is
    # forked after it). Every launch choice and the seeds' margin and slope are chosen by timing the tuning pairs
    # (drawn like the workload, another seed) and frozen before the workload's pairs are timed.
    train_r = train_q = tune_r = tune_q = None
    if not args.plain_only:
        t0 = time.time()
        tune_n = args.tune_pairs or pairs
        tune_r, tune_q = synthetic.uneven_parallel(tune_n, seed=44, workers=workers)
        if train_n > 0:
            train_r, train_q = synthetic.uneven_parallel(train_n, seed=41, workers=workers)
        print(f"[bench] tuning ({tune_n}) and training ({train_n}) pairs ready in {time.time() - t0:.1f}s",
              file=sys.stderr, flush=True)
    # Plain doubling is also timed in a child process at HIP's default 4 hardware queues (more queues let the
    # bins run at once, which helps the other variants and can slow this one): the baseline keeps its better time.
    child = None
    if not args.plain_only and os.environ.get("GPU_MAX_HW_QUEUES") != "4":
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "plain.npz")
            cmd = [sys.executable, os.path.abspath(__file__), "--workload", "uneven", "--pairs", str(pairs), "--steps",
                   str(args.steps), "--warmup", str(args.warmup), "--plain-only", path]
            env = dict(os.environ, DINARA_HW_QUEUES="4")
            if subprocess.run(cmd, env=env, timeout=600).returncode == 0:
                with np.load(path, allow_pickle=False) as z:
                    blob, ends = z["cigars"].tobytes(), z["ends"]
                    starts = np.concatenate(([0], ends[:-1]))
                    child = {"ms": float(z["ms"]), "dev": float(z["dev"]), "apart": z["apart"].tolist(),
                             "results": (z["costs"], [blob[a:b] for a, b in zip(starts, ends)])}
    handle = _lib.lib()
    if handle.dinara_device_count() < 1:
        print("[bench] no GPU visible", file=sys.stderr)
        return 1
    A = _lib.BATCH_ALIGNMENTS
    results = {}

    # Plain band doubling: every band from 64 diagonals keeps its decisions, twice as wide until proven.
    plain, keep_plain = edit_batch(handle, refs, qrys, A, workers)
    handle.dinara_batch_set_retry(plain, 1)
    handle.dinara_batch_set_keep_first(plain, 1)
    handle.dinara_batch_set_persistent(plain, 0)
    handle.dinara_batch_set_graph(plain, 0)
    # Its bins side by side on their streams or one after another: timed, the faster kept (the stronger baseline).
    plain_apart = {}
    for apart in (1, 0):
        handle.dinara_batch_set_side_by_side(plain, apart)
        plain_apart[apart] = timed_steps(handle, plain, 3, 1)[0]
    handle.dinara_batch_set_side_by_side(plain, min(plain_apart, key=plain_apart.get))
    plain_ms, plain_dev = timed_steps(handle, plain, args.steps, args.warmup)
    results["plain"] = edit_results(handle, plain, A, pairs)
    plain_prof = bin_profile(handle, plain)
    handle.dinara_batch_free(plain)
    del keep_plain
    if args.plain_only:
        costs, cigars = results["plain"]
        ends = np.cumsum([len(c) for c in cigars]).astype(np.int64)
        np.savez(args.plain_only, ms=plain_ms, dev=plain_dev, apart=np.array([plain_apart[1], plain_apart[0]]),
                 costs=costs, cigars=np.frombuffer(b"".join(cigars), dtype=np.uint8), ends=ends)
        print(f"[bench] plain doubling at {os.environ.get('GPU_MAX_HW_QUEUES')} hardware queues {plain_ms:.1f} "
              f"ms/step", file=sys.stderr, flush=True)
        return 0
    plain_by_queues = {os.environ.get("GPU_MAX_HW_QUEUES"): round(plain_ms, 3)}
    if child is not None:
        results["plain_4_queues"] = child["results"]
        plain_by_queues["4"] = round(child["ms"], 3)
        if child["ms"] < plain_ms:
            plain_ms, plain_dev = child["ms"], child["dev"]
    print(f"[bench] plain doubling {plain_ms:.1f} ms/step ({plain_by_queues} by hardware queues)", file=sys.stderr,
          flush=True)

    # Every choice below is made on the held-out tuning pairs and frozen; the workload's pairs are then timed once
    # with those choices.
    # Persistent, captured: the autotuner's launch shape (workgroups a compute unit, start bins, bins side by side).
    tb, tkeep = edit_batch(handle, tune_r, tune_q, A, workers)
    chosen = (ctypes.c_int64 * 7)()
    t0 = time.time()
    if handle.dinara_batch_autotune(tb, chosen) != 0:
        raise RuntimeError("autotune failed")
    tune_s = time.time() - t0
    print(f"[bench] persistent shape tuned on held-out pairs in {tune_s:.1f}s: {list(chosen)}", file=sys.stderr,
          flush=True)

    # Learned seeds: a network trained on the aligner's own distances of separate pairs.
    net = None
    learned_info = None
    if train_n > 0:
        t0 = time.time()
        lb, lkeep = edit_batch(handle, train_r, train_q, _lib.BATCH_DISTANCES, workers)
        if handle.dinara_batch_run(lb) != 0:
            raise RuntimeError("labelling run failed")
        labels, _ = edit_results(handle, lb, _lib.BATCH_DISTANCES, train_n)
        x_train = np.zeros((train_n, 288), dtype=np.float32)
        if handle.dinara_batch_features(lb, x_train.ctypes.data) != 0:
            raise RuntimeError("features failed")
        handle.dinara_batch_free(lb)
        del lkeep, train_r, train_q
        y_train = np.log1p(np.maximum(labels, 0)).astype(np.float32)
        net = handle.dinara_mlp_new(0, 17)
        loss = ctypes.c_double(0)
        epochs = args.epochs or 3
        t1 = time.time()
        if handle.dinara_mlp_train(net, train_n, x_train.ctypes.data, y_train.ctypes.data, epochs, 4096, 2e-3, 7,
                                   ctypes.byref(loss)) != 0:
            raise RuntimeError("training failed")
        train_s = time.time() - t1
        label_train_s = time.time() - t0
        del x_train

        def seeder(b, margin, slope):
            return lambda: handle.dinara_batch_seed_learned_by_length(b, net, ctypes.c_double(margin),
                                                                      ctypes.c_double(slope))

        # On the tuning pairs: sweeps keep their decisions (a proven first band is walked at once); graph or not
        # first (the two forms schedule the bins differently, so the choices after it are timed in the chosen one),
        # then the launch shape tuned again for the seeds, margin and slope, and first-bin floors, each timed.
        margin, slope = args.margin, args.slope
        handle.dinara_batch_set_tuning(tb, chosen[0], (ctypes.c_int64 * 5)(0, 0, 0, 0, 0))
        handle.dinara_batch_set_keep_first(tb, 1)
        by_launch = {}
        for graph in (1, 0):
            handle.dinara_batch_set_graph(tb, graph)
            by_launch[graph] = timed_steps(handle, tb, 3, 1, before=seeder(tb, margin, slope))[0]
            print(f"[bench] held-out: learned {'graph' if graph else 'launches one by one'}: {by_launch[graph]:.1f} ms",
                  file=sys.stderr, flush=True)
        use_graph = min(by_launch, key=by_launch.get)
        handle.dinara_batch_set_graph(tb, use_graph)
        seeder(tb, margin, slope)()
        learned_chosen = (ctypes.c_int64 * 7)()
        if handle.dinara_batch_autotune(tb, learned_chosen) != 0:
            raise RuntimeError("autotune failed")
        by_slope = {}
        if args.slopes or args.margins:
            slopes = [float(v) for v in args.slopes.split(",")] if args.slopes else [args.slope]
            margins = [float(v) for v in args.margins.split(",")] if args.margins else [args.margin]
            for mg in margins:
                for sl in slopes:
                    by_slope[(mg, sl)] = timed_steps(handle, tb, 3, 1, before=seeder(tb, mg, sl))[0]
                    print(f"[bench] held-out: margin {mg} slope {sl}: {by_slope[(mg, sl)]:.1f} ms", file=sys.stderr,
                          flush=True)
            margin, slope = min(by_slope, key=by_slope.get)
        by_floor, by_bins = {}, {}
        floor = (0, 0)
        if args.floors:
            for item in args.floors.split(","):
                f3, f4 = (int(v) for v in item.split(":"))
                handle.dinara_batch_set_tuning(tb, learned_chosen[0], (ctypes.c_int64 * 5)(0, 0, 0, f3, f4))
                by_floor[(f3, f4)] = timed_steps(handle, tb, 2, 1, before=seeder(tb, margin, slope))[0]
            floor = min(by_floor, key=by_floor.get)
            handle.dinara_batch_set_tuning(tb, learned_chosen[0], (ctypes.c_int64 * 5)(0, 0, 0, *floor))
        # The first pass's bin edges: a few bin sets, the bins between the powers of two in or out.
        widths = [int(handle.dinara_batch_stat(tb, 1200 + k)) for k in range(int(handle.dinara_batch_stat(tb, 12)))]
        every = (1 << len(widths)) - 1
        bin_sets = {"every bin": every,
                    "powers of two": sum(1 << k for k, w in enumerate(widths) if w & (w - 1) == 0),
                    "every bin but 1344": every & ~(1 << widths.index(1344)),
                    "every bin but 192 and 1344": every & ~(1 << widths.index(192)) & ~(1 << widths.index(1344))}
        for name, mask in bin_sets.items():
            if handle.dinara_batch_set_bins(tb, mask) != 0:
                raise RuntimeError("bin set refused")
            by_bins[name] = timed_steps(handle, tb, 3, 1, before=seeder(tb, margin, slope))[0]
            print(f"[bench] held-out: first bins {name}: {by_bins[name]:.1f} ms", file=sys.stderr, flush=True)
        bin_set = min(by_bins, key=by_bins.get)
        first_bins = bin_sets[bin_set]
    handle.dinara_batch_free(tb)
    del tkeep, tune_r, tune_q

    # The workload's pairs, the persistent path with the frozen shape.
    batch, keep = edit_batch(handle, refs, qrys, A, workers)
    handle.dinara_batch_set_tuning(batch, chosen[0], (ctypes.c_int64 * 5)(*[chosen[i] for i in range(1, 6)]))
    handle.dinara_batch_set_side_by_side(batch, chosen[6])
    persistent_ms, persistent_dev = timed_steps(handle, batch, args.steps, args.warmup)
    results["persistent"] = edit_results(handle, batch, A, pairs)
    prof = bin_profile(handle, batch)
    first_proven = handle.dinara_batch_stat(batch, 4)
    print(f"[bench] persistent {persistent_ms:.1f} ms/step", file=sys.stderr, flush=True)

    # The learned path with the frozen choices: seeding (features, prediction, first bands) inside each step.
    learned_ms = learned_dev = None
    end_to_end = None
    if net is not None:
        handle.dinara_batch_set_tuning(batch, learned_chosen[0], (ctypes.c_int64 * 5)(0, 0, 0, *floor))
        handle.dinara_batch_set_side_by_side(batch, learned_chosen[6])
        handle.dinara_batch_set_graph(batch, use_graph)
        handle.dinara_batch_set_bins(batch, first_bins)
        seed = seeder(batch, margin, slope)
        seed()
        two_pass_ms, _ = timed_steps(handle, batch, args.steps, args.warmup, before=seed)
        results["learned_two_pass"] = edit_results(handle, batch, A, pairs)
        handle.dinara_batch_set_keep_first(batch, 1)
        learned_ms, learned_dev = timed_steps(handle, batch, args.steps, args.warmup, before=seed)
        results["learned"] = edit_results(handle, batch, A, pairs)
        learned_proven = handle.dinara_batch_stat(batch, 4)
        walked_kept = handle.dinara_batch_stat(batch, 9)
        learned_prof = bin_profile(handle, batch)
        handle.dinara_batch_free(batch)
        del keep
        # End to end: packing and upload, seeding, the run, and the alignments copied out (CIGARs malloc'ed).
        end_to_end = end_to_end_learned(handle, refs, qrys, workers, net, margin, slope, learned_chosen, floor,
                                        use_graph, first_bins)
        learned_info = {"train_pairs": train_n, "epochs": epochs, "loss": round(loss.value, 4),
                        "label_and_train_s": round(label_train_s, 1), "train_s": round(train_s, 1),
                        "chosen_on": f"{args.tune_pairs or pairs} held-out pairs drawn like the workload (seed 44)",
                        "margin": margin, "slope": slope,
                        "by_margin_slope_ms_held_out": {f"{k[0]},{k[1]}": round(v, 3) for k, v in by_slope.items()},
                        "first_bin_floors": list(floor), "graph": int(use_graph),
                        "by_launch_ms_held_out": {("graph" if k else "one_by_one"): round(v, 3)
                                                  for k, v in by_launch.items()},
                        "by_floors_ms_held_out": {f"{k[0]}:{k[1]}": round(v, 3) for k, v in by_floor.items()},
                        "first_pass_bins": [w for k, w in enumerate(widths) if (first_bins >> k) & 1],
                        "first_pass_bin_set": bin_set,
                        "by_bin_set_ms_held_out": {k: round(v, 3) for k, v in by_bins.items()},
                        "first_band_proved": round(learned_proven / pairs, 4),
                        "proved_by_a_sweep_keeping_decisions": round(walked_kept / pairs, 4),
                        "two_pass_ms_per_step": round(two_pass_ms, 3), "bins": learned_prof,
                        "tuning": {"grid_per_cu": int(learned_chosen[0]), "side_by_side": int(learned_chosen[6])}}
        handle.dinara_mlp_free(net)
        print(f"[bench] learned {learned_ms:.1f} ms/step", file=sys.stderr, flush=True)
    else:
        handle.dinara_batch_free(batch)
        del keep

    launch_form = ("one captured hipGraph a run" if use_graph else "launches one by one") if net is not None \
        else "one captured hipGraph a run"
    costs, cigars = results["persistent"]
    same_variants = all(np.array_equal(r[0], costs) and r[1] == cigars for r in results.values())
    best_name = min((k for k in ("persistent", "learned") if k in results),
                    key=lambda k: persistent_ms if k == "persistent" else learned_ms)
    value_ms = learned_ms if learned_ms is not None else persistent_ms
    value = pairs / (value_ms / 1e3)

    # Roofline of the dominant kernel, the sweeps of the headline (learned) path: each pair's proven band at its
    # true cost, BIT_OPS_TRACE int32 ops a 64-cell word (the step with its decisions kept), over the sweeps' elapsed
    # time (the bins overlap on their streams: the run launched kernel by kernel, less its walk).
    proven = proven_band_cells(n, m, np.maximum(costs, 0), True, opening=0, extension=1)
    if learned_info is not None:
        sweeps_ms = learned_info["bins"]["run_ms"] - learned_info["bins"]["walk_ms"]
        kernel = "edit_sweep<keep decisions> (every bin, side by side), learned seeds"
    else:
        sweeps_ms = prof["run_ms"] - prof["walk_ms"]
        kernel = "edit_sweep (first pass and tracing pass, every bin)"
    achieved = float(proven.sum()) * BIT_OPS_FIRST / 64 / (sweeps_ms / 1e3) / 1e12
    achieved_trace = float(proven.sum()) * BIT_OPS_TRACE / 64 / (sweeps_ms / 1e3) / 1e12
    peaks = measured_peaks()
    traffic = MEASURED_TRAFFIC.get(("uneven", pairs, "edit_sweep"))
    roofline = {"bound": "valu", "kernel": kernel, "achieved": round(achieved, 3), "peak": VALU_PEAK_TOPS,
                "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 4),
                "traffic": round(traffic["read"] + traffic["written"]) if traffic else None,
                "launch_ms": round(sweeps_ms, 3),
                "achieved_with_decisions": round(achieved_trace, 3),
                "frac_with_decisions": round(achieved_trace / VALU_PEAK_TOPS, 4),
                "measured_peak": peaks["valu_int32_tops"],
                "frac_of_measured_peak": round(achieved / peaks["valu_int32_tops"], 4) if peaks["valu_int32_tops"]
                else None,
                "note": "proven-band cells at the true cost (strict) x 34 int32 ops a 64-cell word (the bit-parallel "
                        "step); _with_decisions counts 56 (the step and its 2-bit decisions); the sweeps' elapsed "
                        "time in a run launched kernel by kernel, less the walk; traffic: HBM bytes the learned run's "
                        "sweeps moved, rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (calibrated on a plain copy)"}
    # The same sweeps against HBM: the 2-bit decisions the walk needs (proven-band cells, algorithmic), and those
    # the sweeps write (every word-column swept keeping decisions, 16 bytes each).
    bins_now = learned_info["bins"] if learned_info is not None else prof
    swept = float(sum(bins_now["sweep_word_columns"]) + sum(bins_now["trace_word_columns"])) * 16
    needed = float(proven.sum()) / 4
    roofline["hbm_writes"] = {"bound": "hbm", "achieved": round(needed / (sweeps_ms / 1e3) / 1e12, 3),
                              "written": round(swept / (sweeps_ms / 1e3) / 1e12, 3), "peak": 8.0, "unit": "TB/s",
                              "frac": round(needed / (sweeps_ms / 1e3) / 8e12, 4),
                              "note": "decision bytes: proven-band cells x 2 bits (achieved) and word-columns swept "
                                      "keeping decisions x 16 B (written), over the same time"}

    cpu = None
    exact_cpu = None
    if not args.no_cpu:
        lib, flags = load_edit_baseline()
        rng = np.random.default_rng(11)
        sample = np.sort(rng.choice(pairs, size=min(pairs, args.cpu_sample), replace=False))
        s_refs = [refs[i] for i in sample]
        s_qrys = [qrys[i] for i in sample]
        R, Q = _lib.Texts(s_refs), _lib.Texts(s_qrys)
        stride = int((n[sample] + m[sample]).max()) * 2 + 2
        cpu_costs = (ctypes.c_int64 * len(sample))()
        buf = ctypes.create_string_buffer(len(sample) * stride)
        t0 = time.perf_counter()
        lib.baseline_edit(len(sample), ctypes.cast(R.pointers, ctypes.c_void_p), ctypes.cast(R.lengths, ctypes.c_void_p),
                          ctypes.cast(Q.pointers, ctypes.c_void_p), ctypes.cast(Q.lengths, ctypes.c_void_p), cores, 0,
                          cpu_costs, buf, stride)
        cpu_s = time.perf_counter() - t0
        raw = buf.raw
        same = sum(int(cpu_costs[k] == costs[i] and raw[k * stride:(k + 1) * stride].split(b"\0", 1)[0] == cigars[i])
                   for k, i in enumerate(sample))
        exact_cpu = same / len(sample)
        cpu = {"value": round(len(sample) / cpu_s, 1), "unit": "pairs/s", "cores": cores, "kind": "port",
               "sample": f"{len(sample)} pairs drawn uniformly from the workload, per-pair bit-parallel band "
                         f"doubling without seeds (slower than the reference's A*PA2), {flags}, {cpu_model()}"}
        print(f"[bench] CPU baseline {len(sample) / cpu_s:.0f} pairs/s on {cores} cores; equal on {same}/{len(sample)}",
              file=sys.stderr, flush=True)

    line = {
        "metric": "pairs/s + GCUPS on 1 MI355X: learned-band vs plain band doubling, bit-exact",
        "value": round(value, 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "unit-cost distance + Ties.LEFT CIGAR, 200,000 pairs log-normal around 1.5 kbp (sigma 1, "
                        "100 bp..20 kbp), 1-15% error (50/25/25), learned band seeds + persistent length-binned "
                        f"launch ({launch_form}), launch shape, margin, slope and first-pass bin edges chosen on "
                        "held-out pairs",
            "pairs": pairs,
        },
        "gcups": round(float((n * m).sum()) / (value_ms / 1e3) / 1e9, 2),
        "variants_ms_per_step": {"plain_doubling": round(plain_ms, 3), "persistent_tuned": round(persistent_ms, 3),
                                 "learned_persistent": round(learned_ms, 3) if learned_ms else None},
        "speedup_vs_plain_doubling": round(plain_ms / value_ms, 3),
        "fastest_variant": best_name,
        "device_ms_per_step": {"plain_doubling": round(plain_dev, 3), "persistent_tuned": round(persistent_dev, 3),
                               "learned_persistent": round(learned_dev, 3) if learned_dev else None},
        "bit_exact_across_variants": bool(same_variants),
        "bit_exact_vs_cpu_sample": exact_cpu,
        "first_band_proved": round(first_proven / pairs, 4),
        "tuning": {"grid_per_cu": int(chosen[0]), "start_bins": [int(chosen[i]) for i in range(1, 6)],
                   "side_by_side": int(chosen[6]), "tune_s": round(tune_s, 1)},
        "learned": learned_info,
        "bins": {"persistent": prof, "plain_doubling": plain_prof},
        "plain_doubling_side_by_side_ms": {("side_by_side" if k else "one_after_another"): round(v, 3)
                                           for k, v in plain_apart.items()},
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
        "plain_doubling_ms_by_hw_queues": plain_by_queues,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "end_to_end": end_to_end,
        "measured_peaks": peaks,
    }
    print(json.dumps(line))
    return 0 if same_variants else 1



def run_edit(args) -> int:
    """Unit-cost distances of 10,000 pairs of 150 bp at 2% (configuration 1's workload): the persistent path, one
    captured graph a run, against the restated lanes CPU path (oracle/baseline.cpp at costs (1, 0, 1))."""
    pairs = args.pairs or 10_000
    refs, qrys = synthetic.edit_short(pairs)
    n = np.fromiter((len(r) for r in refs), dtype=np.int64, count=pairs)
    m = np.fromiter((len(q) for q in qrys), dtype=np.int64, count=pairs)
    handle = _lib.lib()
    if handle.dinara_device_count() < 1:
        print("[bench] no GPU visible", file=sys.stderr)
        return 1
    cores = host_cores()
    D = _lib.BATCH_DISTANCES
    # The bins side by side on streams of their own, or one after another, chosen on held-out pairs drawn alike:
    # with every pair proven by its first band, the empty retry launches cost least in one stream.
    tune_r, tune_q = synthetic.edit_short(pairs, seed=5)
    tb, tkeep = edit_batch(handle, tune_r, tune_q, D, min(cores, 16))
    by_shape = {}
    for side_by_side in (1, 0):
        handle.dinara_batch_set_side_by_side(tb, side_by_side)
        by_shape[side_by_side] = timed_steps(handle, tb, 20, 5)[0]
    handle.dinara_batch_free(tb)
    del tkeep, tune_r, tune_q
    side_by_side = min(by_shape, key=by_shape.get)
    batch, keep = edit_batch(handle, refs, qrys, D, min(cores, 16))
    handle.dinara_batch_set_side_by_side(batch, side_by_side)
    ms, dev = timed_steps(handle, batch, args.steps, args.warmup)
    costs, _ = edit_results(handle, batch, D, pairs)
    prof = bin_profile(handle, batch)
    handle.dinara_batch_free(batch)
    unit = _lib.CCosts(1, 0, 1, -1, 0, 0, 0, -1, 0)
    end_to_end = drop_in_call(handle, *keep, pairs, ctypes.byref(unit), None, min(cores, 16), reps=10,
                              kind=_lib.BATCH_DISTANCES)
    proven = proven_band_cells(n, m, costs, False, opening=0, extension=1)
    first_ms = float(np.sum(prof["sweep_ms"]))
    achieved = float(proven.sum()) * BIT_OPS_FIRST / 64 / (first_ms / 1e3) / 1e12
    roofline = {"bound": "valu", "kernel": "edit_sweep (every bin)", "achieved": round(achieved, 4),
                "peak": VALU_PEAK_TOPS, "unit": "TOP/s", "frac": round(achieved / VALU_PEAK_TOPS, 5),
                "traffic": None, "launch_ms": round(first_ms, 4),
                "note": "proven-band cells x 34 int32 ops a 64-cell word; 10,000 short pairs fill few lanes, so the "
                        "launch is latency-bound"}
    cpu = exact = None
    if not args.no_cpu:
        lib, flags = load_baseline(cores)
        R, Q = keep
        cpu_costs = (ctypes.c_int64 * pairs)()
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            lib.baseline_alignments(pairs, ctypes.cast(R.pointers, ctypes.c_void_p),
                                    ctypes.cast(R.lengths, ctypes.c_void_p), ctypes.cast(Q.pointers, ctypes.c_void_p),
                                    ctypes.cast(Q.lengths, ctypes.c_void_p), 1, 0, 1, cores, cpu_costs, None, 0)
        cpu_s = (time.perf_counter() - t0) / reps
        exact = float(np.mean(np.array(cpu_costs[:]) == costs))
        cpu = {"value": round(pairs / cpu_s, 1), "unit": "pairs/s", "cores": cores, "kind": "port",
               "sample": f"the full workload x {reps}, the lanes path at costs (1, 0, 1), {flags}, {cpu_model()}"}
    line = {
        "metric": "pairs/s (global unit-cost distances, bit-exact)",
        "value": round(pairs / (ms / 1e3), 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "global unit-cost distances, 10,000 pairs of 150 bp at 2% (equal sub/ins/del), bins "
                               f"{'side by side' if side_by_side else 'one after another'} (chosen on held-out pairs)",
                   "pairs": pairs},
        "by_launch_shape_ms_held_out": {("side_by_side" if k else "one_after_another"): round(v, 4)
                                        for k, v in by_shape.items()},
        "gcups": round(float((n * m).sum()) / (ms / 1e3) / 1e9, 2),
        "device_ms_per_step": round(dev, 4),
        "bit_exact_vs_cpu": exact,
        "bins": prof,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "end_to_end": end_to_end,
    }
    print(json.dumps(line))
    return 0


def load_probe():
    path = os.path.join(ROOT, "dinara-align_amd", "probe", "libdinara_probe.so")
    lib = ctypes.CDLL(path)
    lib.probe_valu.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]
    lib.probe_valu.restype = ctypes.c_double
    lib.probe_copy.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]
    lib.probe_copy.restype = ctypes.c_double
    return lib


def measured_peaks():
    """The denominators measured on this GPU: int32 VALU ops a second (8 independent add/xor chains a lane, 8
    workgroups of 256 a compute unit) and a plain 1 GiB device copy's HBM rate (read plus write)."""
    lib = load_probe()
    ops = ctypes.c_double(0)
    valu = lib.probe_valu(8, 4096, 5, ctypes.byref(ops))
    copy_ms = ctypes.c_double(0)
    copy = lib.probe_copy(1 << 30, 10, ctypes.byref(copy_ms))
    return {"valu_int32_tops": round(valu, 2), "hbm_copy_gbs": round(copy, 1),
            "note": "probe/probe.hip: int32 add/xor chains on every compute unit; a 1 GiB copy kernel, 16 B a lane"}


def train_network(handle, refs, qrys, workers, epochs, net_file=""):
    """A network trained on the aligner's own unit-cost distances of `refs`/`qrys` (log(1 + cost) targets), or the
    weights kept in `net_file` (an .npy of float32) when it exists; written there after training when named."""
    net = handle.dinara_mlp_new(0, 17)
    count = handle.dinara_mlp_parameter_count(net)
    if net_file and os.path.exists(net_file):
        weights = np.load(net_file, allow_pickle=False).astype(np.float32)
        if weights.size != count or handle.dinara_mlp_set_weights(net, weights.ctypes.data) != 0:
            raise RuntimeError(f"{net_file} does not hold this network's weights")
        return net, None
    b, keep = edit_batch(handle, refs, qrys, _lib.BATCH_DISTANCES, workers)
    if handle.dinara_batch_run(b) != 0:
        raise RuntimeError("labelling run failed")
    labels, _ = edit_results(handle, b, _lib.BATCH_DISTANCES, len(refs))
    x = np.zeros((len(refs), 288), dtype=np.float32)
    if handle.dinara_batch_features(b, x.ctypes.data) != 0:
        raise RuntimeError("features failed")
    handle.dinara_batch_free(b)
    del keep
    y = np.log1p(np.maximum(labels, 0)).astype(np.float32)
    loss = ctypes.c_double(0)
    if handle.dinara_mlp_train(net, len(refs), x.ctypes.data, y.ctypes.data, epochs, 4096, 2e-3, 7,
                               ctypes.byref(loss)) != 0:
        raise RuntimeError("training failed")
    if net_file:
        weights = np.zeros(count, dtype=np.float32)
        handle.dinara_mlp_get_weights(net, weights.ctypes.data)
        np.save(net_file, weights, allow_pickle=False)
    return net, loss.value


def run_learned_profile(args) -> int:
    """The learned path's run alone, for profilers: the mixed-length workload's 200,000 pairs seeded by the network
    (trained as the default bench trains it, or loaded from --net-file), sweeps keeping their decisions, launches
    one by one so each bin's sweeps are kernels of their own in a trace; `--steps` runs after the training. A plain
    1 GiB device copy (probe/probe.hip) runs first: the copy the HBM byte counters are calibrated against. Prints the
    library's own per-bin times of the last run as one JSON line."""
    pairs = args.pairs or 200_000
    workers = min(host_cores(), 16)
    refs, qrys = synthetic.uneven_parallel(pairs, seed=3, workers=workers)
    train_r = train_q = None
    if not (args.net_file and os.path.exists(args.net_file)):
        train_r, train_q = synthetic.uneven_parallel(args.train or 1_000_000, seed=41, workers=workers)
    print("[profile] pairs ready", file=sys.stderr, flush=True)
    handle = _lib.lib()
    copy_ms = ctypes.c_double(0)
    copy = load_probe().probe_copy(1 << 30, 2, ctypes.byref(copy_ms))
    print(f"[profile] calibration copy: 1 GiB in {copy_ms.value:.3f} ms ({copy:.0f} GB/s)", file=sys.stderr,
          flush=True)
    net, loss = train_network(handle, train_r, train_q, workers, args.epochs or 3, args.net_file)
    del train_r, train_q
    print(f"[profile] network ready (loss {loss})", file=sys.stderr, flush=True)
    b, keep = edit_batch(handle, refs, qrys, _lib.BATCH_ALIGNMENTS, workers)
    handle.dinara_batch_set_tuning(b, args.grid_per_cu, (ctypes.c_int64 * 5)(0, 0, 0, 0, 0))
    handle.dinara_batch_set_keep_first(b, 1)
    handle.dinara_batch_set_graph(b, 0)
    for step in range(args.steps):
        if handle.dinara_batch_seed_learned_by_length(b, net, ctypes.c_double(args.margin),
                                                      ctypes.c_double(args.slope)) != 0 or \
                handle.dinara_batch_run(b) != 0:
            raise RuntimeError("learned run failed")
        print(f"[profile] learned run {step + 1}: {handle.dinara_batch_stat(b, 0):.2f} ms", file=sys.stderr,
              flush=True)
    bins, widths = edit_bins(handle, b)
    prof = {k: [round(handle.dinara_batch_stat(b, base + i), 4) for i in range(bins)]
            for k, base in (("sweep_ms", 600), ("sweeps", 700), ("trace_ms", 800), ("traces", 900),
                            ("sweep_word_columns", 1000), ("trace_word_columns", 1100))}
    prof["widths"] = widths
    prof["walk_ms"] = handle.dinara_batch_stat(b, 5)
    prof["run_ms"] = handle.dinara_batch_stat(b, 0)
    prof["first_band_proved"] = handle.dinara_batch_stat(b, 4) / pairs
    costs, _ = edit_results(handle, b, _lib.BATCH_ALIGNMENTS, pairs)
    handle.dinara_batch_free(b)
    handle.dinara_mlp_free(net)
    n = np.fromiter((len(r) for r in refs), dtype=np.int64, count=pairs)
    m = np.fromiter((len(q) for q in qrys), dtype=np.int64, count=pairs)
    proven = float(proven_band_cells(n, m, np.maximum(costs, 0), True, opening=0, extension=1).sum())
    print(json.dumps({"profile": "learned run, mixed-length workload", "margin": args.margin, "slope": args.slope,
                      "grid_per_cu": args.grid_per_cu, "proven_band_cells": proven,
                      "full_matrix_cells": float((n * m).sum()), "bins": prof}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--pairs", type=int, default=None, help="pairs (default: the workload's own count)")
    ap.add_argument("--workload", choices=("uneven", "affine", "local", "learned", "edit", "oversized"),
                    default="uneven",
                    help="uneven: unit-cost CIGARs of 200k uneven pairs, learned band + persistent launches against "
                         "plain band doubling (the default); affine: global gap-affine (4,6,2) CIGARs, 100k x 250 bp; "
                         "local: local alignment under a 20-letter table, 50k x (3000 aa window, 300 aa query); "
                         "learned: gap-affine distances of 200k uneven pairs, the learned band against plain band "
                         "doubling; edit: unit-cost distances, 10k x 150 bp; oversized: gap-affine CIGARs of 70 kbp "
                         "pairs, past every band kernel's letters (the whole-matrix path)")
    ap.add_argument("--train", type=int, default=None, help="learned: training pairs (default 1,000,000)")
    ap.add_argument("--epochs", type=int, default=None, help="learned: training epochs (default 3)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=20_000, help="uneven: pairs the CPU baseline times")
    ap.add_argument("--margin", type=float, default=0.0,
                    help="uneven: learned seeds e^(prediction + margin + slope ln(max(longer, 1000) / 1000)) - 1")
    ap.add_argument("--slope", type=float, default=0.0, help="uneven: the seeds' margin per e-fold of length")
    ap.add_argument("--slopes", default="0,0.1,0.2",
                    help="uneven: comma-separated slopes to time on the batch, keeping the fastest ('' for --slope)")
    ap.add_argument("--tune-pairs", type=int, default=None,
                    help="uneven: held-out pairs the launch choices, margin and slope are timed on (default: as many "
                         "as the workload)")
    ap.add_argument("--profile-learned", action="store_true",
                    help="run only the learned path's runs of the mixed-length workload (launches one by one), for "
                         "rocprofv3; --margin, --slope and --grid-per-cu fix its choices")
    ap.add_argument("--grid-per-cu", type=int, default=8, help="--profile-learned: persistent workgroups a CU")
    ap.add_argument("--net-file", default="", help="--profile-learned: load the network from this .npy, or train "
                                                      "and save it there")
    ap.add_argument("--plain-only", default="", help=argparse.SUPPRESS)  # the child timing plain doubling
    ap.add_argument("--floors", default="",
                    help="uneven: least first bins (4k-16k letters : 16k and more) timed for the learned seeds, e.g. "
                         "'0:0,0:6,5:6' (off by default: on the uneven batch none beat 0:0)")
    ap.add_argument("--margins", default="-0.1,0,0.1",
                    help="uneven: comma-separated margins timed with each slope ('' for --margin)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (pairs would shard by length-balanced "
              "ranges, one process per GPU); run with --gpus 1", file=sys.stderr)
        return 2
    if args.profile_learned:
        return run_learned_profile(args)
    if args.workload == "local":
        return run_local(args)
    if args.workload == "learned":
        return run_learned(args)
    if args.workload == "uneven":
        return run_uneven(args)
    if args.workload == "edit":
        return run_edit(args)
    if args.workload == "oversized":
        return run_oversized(args)
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
    end_to_end = drop_in_call(handle, R, Q, pairs, ctypes.byref(costs), ctypes.byref(options), min(cores, 16))
    print(f"[bench] drop-in dinara_gpu_alignments end to end: {end_to_end['ms']:.1f} ms "
          f"({end_to_end['pairs_per_s']:.0f} pairs/s)", file=sys.stderr, flush=True)

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
    measured = MEASURED_TRAFFIC.get(("affine", pairs, CLASS_NAMES[dominant]))
    traffic = None
    if measured is not None:
        moved = measured["read"] + measured["written"]
        traffic = {"bytes": moved, "read": measured["read"], "written": measured["written"],
                   "hbm_gbs": round(moved / (launch_ms / 1e3) / 1e9, 1),
                   "hbm_frac": round(moved / (launch_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    roofline = {
        "bound": "valu",
        "kernel": CLASS_NAMES[dominant],
        "achieved": round(achieved, 3),
        "peak": VALU_PEAK_TOPS,
        "unit": "TOP/s",
        "frac": round(achieved / VALU_PEAK_TOPS, 4),
        "traffic": traffic,
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
        "end_to_end": end_to_end,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())

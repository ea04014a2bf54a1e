"""Parity at the benchmark's full sizes, through the C ABI.

The mixed-length workload: 200,000 pairs, lengths log-normal around 1.5 kbp clipped to 100 bp..20 kbp, 1-15%
error (the benchmark's own pairs, seed 3).
- Unit costs, Ties.LEFT CIGARs: plain band doubling, the persistent length-binned path and learned band seeds
  (sweeps keeping their decisions) give the same costs and CIGARs byte for byte; every CIGAR spells its pair and
  prices at its cost; the costs equal a distances batch's; the 20 longest pairs and 300 drawn at random equal the
  oracle's (its banded-memory form, which stores only the strict proven band).
- Gap-affine (4, 6, 2) distances of 200,000 fresh pairs (seed 42): certified retries, plain doubling and learned
  seeds give the same costs; the 20 longest and 300 random pairs equal the oracle's.
"""

import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from dinara_align import _lib, synthetic
from oracle import oracle as o

pytestmark = pytest.mark.gpu

PAIRS = 200_000
WORKERS = min(16, len(os.sched_getaffinity(0)))
UNIT = (1, 0, 1)
AFFINE = (4, 6, 2)


def say(text):
    print(f"[full] {text}", flush=True)


class Resident:
    """A resident batch (dinara_batch_new) under one-piece costs."""

    def __init__(self, refs, qrys, kind, costs):
        self.h = _lib.lib()
        self.R, self.Q = _lib.Texts(refs), _lib.Texts(qrys)
        self.costs = _lib.CCosts(*costs, -1, 0, 0, 0, -1, 0)
        self.n = len(refs)
        self.kind = kind
        status = ctypes.c_int64(0)
        self.ptr = self.h.dinara_batch_new(0, kind, self.n, *self.R.args(), *self.Q.args(), ctypes.byref(self.costs),
                                           None, None, WORKERS, ctypes.byref(status))
        assert self.ptr, status.value

    def run(self):
        assert self.h.dinara_batch_run(self.ptr) == 0
        if self.kind == _lib.BATCH_DISTANCES:
            out = np.zeros(self.n, dtype=np.int64)
            assert self.h.dinara_batch_distances(self.ptr, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))) == 0
            return out
        found = (_lib.CAlignment * self.n)()
        statuses = (ctypes.c_int64 * self.n)()
        assert self.h.dinara_batch_alignments(self.ptr, found, statuses) == 0
        costs = np.zeros(self.n, dtype=np.int64)
        cigars = []
        for i in range(self.n):
            assert statuses[i] == 0, (i, statuses[i])
            costs[i] = found[i].cost
            cigars.append(ctypes.string_at(found[i].cigar, found[i].cigar_length))
            self.h.dinara_free(found[i].cigar)
        return costs, cigars

    def features(self):
        out = np.zeros((self.n, 288), dtype=np.float32)
        assert self.h.dinara_batch_features(self.ptr, out.ctypes.data) == 0
        return out

    def free(self):
        self.h.dinara_batch_free(self.ptr)


def trained_net(costs, seed):
    """A network trained on the aligner's own distances of 40,000 separate pairs (two epochs)."""
    refs, qrys = synthetic.uneven_parallel(40_000, seed=seed, workers=WORKERS)
    b = Resident(refs, qrys, _lib.BATCH_DISTANCES, costs)
    labels = b.run()
    x = b.features()
    b.free()
    y = np.log1p(np.maximum(labels, 0)).astype(np.float32)
    h = _lib.lib()
    net = h.dinara_mlp_new(0, 17)
    assert net
    loss = ctypes.c_double(0)
    assert h.dinara_mlp_train(net, len(y), x.ctypes.data, y.ctypes.data, 2, 4096, 2e-3, 7, ctypes.byref(loss)) == 0
    say(f"network trained on 40,000 pairs, loss {loss.value:.3f}")
    return net


def sample(refs, qrys, count, seed):
    longest = np.argsort(-np.array([max(len(r), len(q)) for r, q in zip(refs, qrys)]), kind="stable")[:20]
    rng = np.random.default_rng(seed)
    drawn = rng.choice(len(refs), size=count, replace=False)
    return sorted(set(longest.tolist()) | set(drawn.tolist()))


@pytest.fixture(scope="module")
def mixed():
    refs, qrys = synthetic.uneven_parallel(PAIRS, seed=3, workers=WORKERS)
    say(f"{len(refs)} mixed-length pairs ready")
    return refs, qrys


def test_mixed_unit_cost_cigars_at_full_size(mixed):
    refs, qrys = mixed
    h = _lib.lib()
    A = _lib.BATCH_ALIGNMENTS
    results = {}
    plain = Resident(refs, qrys, A, UNIT)
    assert h.dinara_batch_set_retry(plain.ptr, 1) == 0
    assert h.dinara_batch_set_keep_first(plain.ptr, 1) == 0
    assert h.dinara_batch_set_persistent(plain.ptr, 0) == 0
    assert h.dinara_batch_set_graph(plain.ptr, 0) == 0
    results["plain doubling"] = plain.run()
    plain.free()
    say("plain doubling done")
    persistent = Resident(refs, qrys, A, UNIT)
    results["persistent"] = persistent.run()
    say("persistent done")
    net = trained_net(UNIT, seed=41)
    try:
        assert h.dinara_batch_set_keep_first(persistent.ptr, 1) == 0
        assert h.dinara_batch_seed_learned_by_length(persistent.ptr, net, ctypes.c_double(0.1),
                                                     ctypes.c_double(0.1)) == 0
        results["learned"] = persistent.run()
        proven_first = h.dinara_batch_stat(persistent.ptr, 4)
    finally:
        h.dinara_mlp_free(net)
        persistent.free()
    say(f"learned done: {proven_first:.0f} pairs proven by their first band")
    costs, cigars = results["persistent"]
    for name, (other_costs, other_cigars) in results.items():
        assert np.array_equal(other_costs, costs), name
        assert other_cigars == cigars, name
    # Every CIGAR spells its pair and prices at its cost; the costs are the distances batch's.
    bad, first = o.check_alignments(refs, qrys, o.Costs.edit(), cigars, costs)
    assert bad == 0, (bad, first)
    distances = Resident(refs, qrys, _lib.BATCH_DISTANCES, UNIT)
    assert np.array_equal(distances.run(), costs)
    distances.free()
    say("CIGARs checked, distances agree")
    # The longest pairs and a random draw against the oracle.
    picked = sample(refs, qrys, 300, seed=5)
    with ThreadPoolExecutor(WORKERS) as pool:
        want = list(pool.map(lambda i: o.costs_align_proven(refs[i], qrys[i], o.Costs.edit()), picked))
    for i, (cost, cigar) in zip(picked, want):
        assert (costs[i], cigars[i].decode()) == (cost, cigar), (i, len(refs[i]), len(qrys[i]))
    say(f"{len(picked)} pairs equal the oracle, the longest {max(len(refs[i]) for i in picked)} letters")


def test_mixed_gap_affine_distances_at_full_size():
    refs, qrys = synthetic.uneven_parallel(PAIRS, seed=42, workers=WORKERS)
    say(f"{len(refs)} fresh mixed-length pairs ready")
    h = _lib.lib()
    D = _lib.BATCH_DISTANCES
    results = {}
    certified = Resident(refs, qrys, D, AFFINE)
    results["certified retries"] = certified.run()
    certified.free()
    doubling = Resident(refs, qrys, D, AFFINE)
    assert h.dinara_batch_set_retry(doubling.ptr, 1) == 0
    results["plain doubling"] = doubling.run()
    doubling.free()
    net = trained_net(AFFINE, seed=43)
    learned = Resident(refs, qrys, D, AFFINE)
    try:
        assert h.dinara_batch_seed_learned_by_length(learned.ptr, net, ctypes.c_double(0.1),
                                                     ctypes.c_double(0.1)) == 0
        results["learned"] = learned.run()
    finally:
        h.dinara_mlp_free(net)
        learned.free()
    base = results["certified retries"]
    for name, other in results.items():
        assert np.array_equal(other, base), name
    say("certified, doubling and learned distances agree")
    picked = sample(refs, qrys, 300, seed=6)
    costs = o.Costs.affine(*AFFINE)
    with ThreadPoolExecutor(WORKERS) as pool:
        want = list(pool.map(lambda i: o.costs_distance_proven(refs[i], qrys[i], costs), picked))
    assert [int(base[i]) for i in picked] == want
    say(f"{len(picked)} pairs equal the oracle")

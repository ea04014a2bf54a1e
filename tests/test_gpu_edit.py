"""GPU parity of the unit-cost path (the banded bit-parallel sweep, edit.hip) through the C ABI: distances and
CIGARs against the oracle bit for bit, on bands of one lane up to 64 lanes of four words; and every launch shape
the batch can take (the captured graph or launches one by one, refilling lane groups or one pair a group, band
doubling or the proven band, a tracing arena too small for the batch, the autotuner's choice) changes no result.

The unit-cost CIGAR is the reference's bit-parallel traceback rule (traceback.mojo:400-575), which at unit costs
equals its DP-flag rule (lanes.mojo:905-923); the oracle follows the flag rule and is pinned against both.
"""

import ctypes
import os
import random

import numpy as np
import pytest

import dinara_align as da
from dinara_align import _lib, synthetic
from oracle import oracle as o

pytestmark = pytest.mark.gpu

EDIT = o.Costs.affine(1, 0, 1)


def mutated(text, rate, rng, alphabet="ACGT"):
    out = []
    for ch in text:
        u = rng.random()
        if u < rate / 2:
            out.append(rng.choice([c for c in alphabet if c != ch]))
        elif u < 3 * rate / 4:
            out.append(rng.choice(alphabet))
            out.append(ch)
        elif u < rate:
            continue
        else:
            out.append(ch)
    return "".join(out)


def pairs_of(count, rng, lengths, rates, alphabet="ACGT"):
    refs, qrys = [], []
    for i in range(count):
        r = "".join(rng.choice(alphabet) for _ in range(rng.randint(*lengths)))
        refs.append(r)
        qrys.append(mutated(r, rates[i % len(rates)], rng, alphabet))
    return refs, qrys


class Batch:
    """A resident batch under unit costs (or `costs`), with the unit path's switches."""

    def __init__(self, refs, qrys, kind, costs=(1, 0, 1), ties_right=False, eqx=True, max_cost=-1):
        self.h = _lib.lib()
        self.R, self.Q = _lib.Texts(refs), _lib.Texts(qrys)
        mismatch, opening, extension = costs
        self.costs = _lib.CCosts(mismatch, opening, extension, -1, 0, 0, 0, -1, 0)
        self.options = _lib.COptions(-(1 << 60), 1 << 60, max_cost, 1 if eqx else 0, 1 if ties_right else 0,
                                     80 << 20)
        status = ctypes.c_int64(0)
        self.n = len(refs)
        self.kind = kind
        self.ptr = self.h.dinara_batch_new(0, kind, self.n, *self.R.args(), *self.Q.args(), ctypes.byref(self.costs),
                                           None, ctypes.byref(self.options), 4, ctypes.byref(status))
        assert self.ptr, status.value

    def stat(self, what):
        return self.h.dinara_batch_stat(self.ptr, what)

    def widths(self):
        """The unit path's bins, their widths in diagonals: 64 a lane, 1 to 64 lanes of one word, then 64 lanes of
        two and four words."""
        return [int(self.stat(1200 + k)) for k in range(int(self.stat(12)))]

    def run(self):
        assert self.h.dinara_batch_run(self.ptr) == 0
        if self.kind == 0:
            out = (ctypes.c_int64 * self.n)()
            assert self.h.dinara_batch_distances(self.ptr, out) == 0
            return list(out)
        found = (_lib.CAlignment * self.n)()
        statuses = (ctypes.c_int64 * self.n)()
        assert self.h.dinara_batch_alignments(self.ptr, found, statuses) == 0
        out = []
        for i in range(self.n):
            if statuses[i] == 0:
                out.append((found[i].cost, ctypes.string_at(found[i].cigar, found[i].cigar_length).decode()))
                self.h.dinara_free(found[i].cigar)
            else:
                out.append(statuses[i])
        return out

    def free(self):
        self.h.dinara_batch_free(self.ptr)


def oracle_alignments(refs, qrys, right=False, eqx=True):
    return [o.costs_align(r, q, EDIT, (o.NO_BAND_LOW, o.NO_BAND_HIGH), right_ties=right, eqx=eqx)
            for r, q in zip(refs, qrys)]


def test_unit_costs_take_the_bit_parallel_path():
    refs, qrys = ["ACGTACGTTTGCA", "ACGT"], ["ACGTCGTTTTGCA", "AGT"]
    b = Batch(refs, qrys, 1)
    try:
        assert b.stat(6) == 1
        assert b.run() == [(2, "4=1D2=1I6="), (1, "1=1D2=")]
        assert b.stat(7) == 1  # the run replayed its captured graph
    finally:
        b.free()
    affine = Batch(refs, qrys, 0, costs=(4, 6, 2))
    try:
        assert affine.stat(6) == 0
    finally:
        affine.free()


@pytest.mark.parametrize("right", [False, True])
def test_bins_of_several_lanes_against_the_oracle(right):
    # 64 to 4096 diagonals: pairs of 300..2500 letters at up to 50% edits take bins of 1..64 lanes.
    rng = random.Random(31 + right)
    refs, qrys = pairs_of(36, rng, (300, 2500), (0.01, 0.05, 0.15, 0.3, 0.5))
    b = Batch(refs, qrys, 1, ties_right=right)
    try:
        got = b.run()
        widths = b.widths()
        swept = [b.stat(700 + k) for k in range(len(widths))]
    finally:
        b.free()
    assert got == oracle_alignments(refs, qrys, right=right)
    assert sum(1 for w, n in zip(widths, swept) if 128 < w <= 4096 and n > 0) >= 3, swept


@pytest.mark.parametrize("right,keep", [(False, 0), (True, 0), (False, 1), (True, 1)])
def test_bins_between_the_powers_of_two(right, keep):
    # Bands of 192, 384, 768 and 1344 diagonals: 3, 6, 12 and 21 lanes a pair, a wave holding 21, 10, 5 and 3 pairs
    # and leaving its last lanes idle. Seeded guesses place pairs of 1,000-1,800 letters in each; under keep_first the
    # sweep that proves a band keeps its decisions, otherwise a second sweep traces it.
    rng = random.Random(71 + 2 * right + keep)
    refs, qrys = pairs_of(48, rng, (1000, 1800), (0.02, 0.05))
    guesses = np.array([150, 300, 600, 1100] * 12, dtype=np.int64)
    h = _lib.lib()
    b = Batch(refs, qrys, 1, ties_right=right)
    try:
        h.dinara_batch_set_keep_first(b.ptr, keep)
        assert h.dinara_batch_seed(b.ptr, guesses.ctypes.data) == 0
        got = b.run()
        widths = b.widths()
        swept = {w: b.stat(700 + k) for k, w in enumerate(widths)}
    finally:
        b.free()
    assert got == oracle_alignments(refs, qrys, right=right)
    assert all(swept[w] > 0 for w in (192, 384, 768, 1344)), swept
    b = Batch(refs, qrys, 0)
    try:
        assert h.dinara_batch_seed(b.ptr, guesses.ctypes.data) == 0
        assert b.run() == [cost for cost, _ in got]
    finally:
        b.free()


@pytest.mark.parametrize("keep", [1, 0])
def test_first_pass_bin_sets(keep):
    # The first pass's bin edges: host seeds above every pair's cost (so every first band proves its pair) that ask
    # for bands of 192, 384, 768 and 1344 diagonals, filed under several bin sets. With the powers of two only,
    # every first band takes 256, 512, 1024 or 2048 diagonals and the bins between sweep nothing; a set with no bin
    # wide enough falls back to the narrowest of all; device seeding follows the set too. No result changes.
    rng = random.Random(91 + keep)
    refs, qrys = pairs_of(48, rng, (1000, 1800), (0.02, 0.05))
    want = oracle_alignments(refs, qrys)
    guesses = np.array([150, 300, 600, 1100] * 12, dtype=np.int64)
    assert all(cost <= guess for (cost, _), guess in zip(want, guesses))
    h = _lib.lib()
    b = Batch(refs, qrys, 1)
    try:
        widths = b.widths()
        powers = sum(1 << k for k, w in enumerate(widths) if w & (w - 1) == 0)
        assert h.dinara_batch_set_bins(b.ptr, 0) == _lib.INVALID_MODE
        assert h.dinara_batch_set_bins(b.ptr, 1 << len(widths)) == _lib.INVALID_MODE
        assert int(b.stat(13)) == (1 << len(widths)) - 1
        h.dinara_batch_set_keep_first(b.ptr, keep)
        for bins in (powers, 1, (1 << len(widths)) - 1, sum(1 << k for k, w in enumerate(widths) if w in (512, 4096))):
            assert h.dinara_batch_set_bins(b.ptr, bins) == 0
            assert int(b.stat(13)) == bins
            assert h.dinara_batch_seed(b.ptr, guesses.ctypes.data) == 0
            assert b.run() == want, bins
            swept = [b.stat(700 + k) for k in range(len(widths))]
            assert sum(swept) == len(refs), (bins, swept)  # every first band proved its pair
            if bins == powers:
                assert all(swept[k] == 0 for k, w in enumerate(widths) if w & (w - 1)), swept
                assert all(swept[widths.index(w)] > 0 for w in (256, 512, 1024, 2048)), swept
        net = h.dinara_mlp_new(0, 5)
        try:
            for bins in (powers, (1 << len(widths)) - 1):
                assert h.dinara_batch_set_bins(b.ptr, bins) == 0
                assert h.dinara_batch_seed_learned(b.ptr, net, ctypes.c_double(0.0)) == 0
                assert b.run() == want, bins
        finally:
            h.dinara_mlp_free(net)
    finally:
        b.free()


def test_words_of_two_and_four_a_lane():
    # Unrelated sequences of 5,000..9,000 letters: distances near half their length need bands of 8192 and 16384
    # diagonals (64 lanes of two and four words).
    rng = random.Random(37)
    refs = ["".join(rng.choice("ACGT") for _ in range(rng.randint(5000, 9000))) for _ in range(4)]
    qrys = ["".join(rng.choice("ACGT") for _ in range(rng.randint(5000, 9000))) for _ in range(4)]
    b = Batch(refs, qrys, 0)
    try:
        got = b.run()
        widths = b.widths()
        swept = [b.stat(700 + k) for k in range(len(widths))]
    finally:
        b.free()
    assert got == [o.edit_distance(r, q) for r, q in zip(refs, qrys)]
    assert sum(n for w, n in zip(widths, swept) if w > 4096) > 0, swept
    # CIGARs through bands of two words a lane: a length difference of 5000 needs more than 4096 diagonals.
    r = refs[0][:6000]
    qs = [mutated(r[1000:2000], 0.1, rng), "".join(rng.choice("ACGT") for _ in range(900))]
    b = Batch([r, r], qs, 1)
    try:
        got = b.run()
        widths = b.widths()
        traced = [b.stat(900 + k) for k in range(len(widths))]
    finally:
        b.free()
    assert got == oracle_alignments([r, r], qs)
    assert sum(n for w, n in zip(widths, traced) if w > 4096) == 2, traced


def test_byte_letters_and_scaled_costs():
    rng = random.Random(41)
    refs, qrys = pairs_of(60, rng, (0, 400), (0.05, 0.3), alphabet="ACDEFGHIKLMNPQRSTVWY")
    b = Batch(refs, qrys, 1)
    try:
        assert b.run() == oracle_alignments(refs, qrys)
    finally:
        b.free()
    # Costs three times unit: every cost three times, the same CIGARs.
    refs, qrys = pairs_of(60, rng, (0, 300), (0.05, 0.3))
    b = Batch(refs, qrys, 1, costs=(3, 0, 3))
    try:
        got = b.run()
    finally:
        b.free()
    want = oracle_alignments(refs, qrys)
    assert got == [(3 * c, cigar) for c, cigar in want]


def test_caps_and_edges():
    rng = random.Random(43)
    refs, qrys = pairs_of(80, rng, (0, 300), (0.0, 0.1, 0.4))
    refs += ["", "A", "", "ACGT" * 30, "A" * 70]
    qrys += ["", "", "C", "", "T" * 3]
    want = [o.edit_distance(r, q) for r, q in zip(refs, qrys)]
    for cap in (0, 7, 40):
        b = Batch(refs, qrys, 0, max_cost=cap)
        try:
            got = b.run()
        finally:
            b.free()
        assert got == [w if w <= cap else -4 for w in want]  # DINARA_ABOVE_MAX
        b = Batch(refs, qrys, 1, max_cost=cap)
        try:
            got = b.run()
        finally:
            b.free()
        oracle = oracle_alignments(refs, qrys)
        assert got == [w if w[0] <= cap else -4 for w in oracle]


def uneven(count, seed):
    return synthetic.uneven(count, seed=seed, clip=(50, 3000))


def results_with(refs, qrys, kind, setup):
    b = Batch(refs, qrys, kind)
    try:
        setup(b)
        first = b.run()
        again = b.run()  # a second run replays what the first captured
        assert first == again
        return first, b
    finally:
        b.free()


def test_launch_shapes_change_no_result():
    refs, qrys = uneven(400, 47)
    want = oracle_alignments(refs, qrys)
    h = _lib.lib()
    shapes = {
        "graph": lambda b: None,
        "eager": lambda b: h.dinara_batch_set_graph(b.ptr, 0),
        "one pair a group": lambda b: h.dinara_batch_set_persistent(b.ptr, 0),
        "doubling": lambda b: h.dinara_batch_set_retry(b.ptr, 1),
        "wide start": lambda b: h.dinara_batch_set_tuning(b.ptr, 2, (ctypes.c_int64 * 5)(3, 3, 4, 5, 6)),
        "keep first": lambda b: h.dinara_batch_set_keep_first(b.ptr, 1),
        "keep first, doubling, one pair a group": lambda b: (h.dinara_batch_set_keep_first(b.ptr, 1),
                                                            h.dinara_batch_set_retry(b.ptr, 1),
                                                            h.dinara_batch_set_persistent(b.ptr, 0)),
        # The learned path's form: the pairs ready when the retries end are walked while the queue counts come back.
        "keep first, launches one by one": lambda b: (h.dinara_batch_set_keep_first(b.ptr, 1),
                                                     h.dinara_batch_set_graph(b.ptr, 0)),
    }
    for name, setup in shapes.items():
        got, _ = results_with(refs, qrys, 1, setup)
        assert got == want, name


@pytest.mark.parametrize("keep", [1, 0])
def test_later_retries_and_a_second_walk(keep):
    # First bands of no cost, widened by doubling: nearly every pair fails its first band and many fail the doubled
    # one, so later retries run (a retry by the certified band proves its pair at once). Launched one by one, a run
    # walks the pairs ready when the retries end while the queue counts come back, then the later retries (and,
    # without keep_first, the tracing sweeps), then a second walk for what they proved.
    refs, qrys = uneven(300, 83)
    want = oracle_alignments(refs, qrys)
    h = _lib.lib()
    zeros = np.zeros(len(refs), dtype=np.int64)
    b = Batch(refs, qrys, 1)
    try:
        h.dinara_batch_set_keep_first(b.ptr, keep)
        h.dinara_batch_set_graph(b.ptr, 0)
        h.dinara_batch_set_retry(b.ptr, 1)
        assert h.dinara_batch_seed(b.ptr, zeros.ctypes.data) == 0
        got = b.run()
        again = b.run()
        swept = sum(b.stat(700 + k) for k in range(len(b.widths())))
    finally:
        b.free()
    assert got == want and again == want
    assert swept > 2 * len(refs), swept  # some pair was swept three times or more: later retries ran


def test_keep_first_walks_proven_bands_at_once():
    refs, qrys = uneven(300, 67)
    want = oracle_alignments(refs, qrys)
    h = _lib.lib()
    b = Batch(refs, qrys, 1)
    try:
        h.dinara_batch_set_keep_first(b.ptr, 1)
        got = b.run()
        walked = b.stat(9)
        # Seeds at each pair's true cost: every first band proves its pair and is walked as swept.
        seeds = np.ascontiguousarray([c for c, _ in want], dtype=np.int64)
        assert h.dinara_batch_seed(b.ptr, seeds.ctypes.data) == 0
        seeded = b.run()
        seeded_walked = b.stat(9)
    finally:
        b.free()
    assert got == want and seeded == want
    assert 0 < walked <= len(refs)
    assert seeded_walked == len(refs)


def test_keep_first_with_a_small_arena(monkeypatch):
    # Bands that find no room sweep without keeping decisions and are traced in a second sweep.
    refs, qrys = uneven(200, 71)
    monkeypatch.setenv("DINARA_TRACE_ARENA", str(1 << 20))
    h = _lib.lib()
    b = Batch(refs, qrys, 1)
    try:
        h.dinara_batch_set_keep_first(b.ptr, 1)
        got = b.run()
    finally:
        b.free()
    assert got == oracle_alignments(refs, qrys)


def test_small_arena_traces_in_several_passes(monkeypatch):
    refs, qrys = uneven(300, 53)
    want = oracle_alignments(refs, qrys)
    monkeypatch.setenv("DINARA_TRACE_ARENA", str(1 << 20))
    b = Batch(refs, qrys, 1)
    try:
        got = b.run()
        passes = b.stat(8)
        again = b.run()
    finally:
        b.free()
    assert passes >= 2
    assert got == want and again == want


def test_autotune_keeps_every_result():
    refs, qrys = uneven(600, 59)
    h = _lib.lib()
    b = Batch(refs, qrys, 1)
    try:
        chosen = (ctypes.c_int64 * 7)()
        assert h.dinara_batch_autotune(b.ptr, chosen) == 0
        bins = len(b.widths())
        assert chosen[0] in (2, 4, 8, 16) and all(0 <= chosen[i] < bins for i in range(1, 6))
        got = b.run()
    finally:
        b.free()
    assert got == oracle_alignments(refs, qrys)


def test_seeds_change_no_result():
    refs, qrys = uneven(300, 61)
    want = [o.edit_distance(r, q) for r, q in zip(refs, qrys)]
    h = _lib.lib()
    rng = np.random.default_rng(3)
    for guess in (np.array(want) * 2, np.zeros(len(want)), rng.integers(0, 50, len(want))):
        b = Batch(refs, qrys, 0)
        try:
            seeds = np.ascontiguousarray(guess, dtype=np.int64)
            assert h.dinara_batch_seed(b.ptr, seeds.ctypes.data) == 0
            assert b.run() == want
        finally:
            b.free()


def test_learned_seeds_on_the_device_change_no_result():
    # An untrained network (poor seeds) and one trained briefly on the aligner's own distances: the device files
    # every pair's first band from the prediction, and every result still equals the oracle's.
    refs, qrys = uneven(400, 73)
    want = oracle_alignments(refs, qrys)
    h = _lib.lib()
    train_r, train_q = uneven(3000, 79)
    t = Batch(train_r, train_q, 0)
    try:
        x = np.zeros((len(train_r), 288), dtype=np.float32)
        assert h.dinara_batch_features(t.ptr, x.ctypes.data) == 0
        labels = np.log1p(np.array(t.run(), dtype=np.float64)).astype(np.float32)
    finally:
        t.free()
    for trained in (False, True):
        net = h.dinara_mlp_new(0, 9)
        try:
            if trained:
                loss = ctypes.c_double(0)
                assert h.dinara_mlp_train(net, len(labels), x.ctypes.data, labels.ctypes.data, 8, 256, 2e-3, 7,
                                          ctypes.byref(loss)) == 0
            for keep_first in (0, 1):
                b = Batch(refs, qrys, 1)
                try:
                    h.dinara_batch_set_keep_first(b.ptr, keep_first)
                    assert h.dinara_batch_seed_learned(b.ptr, net, ctypes.c_double(0.05)) == 0
                    assert b.run() == want, (trained, keep_first)
                    assert b.run() == want  # the replayed graph reads the same device-filed queues
                finally:
                    b.free()
        finally:
            h.dinara_mlp_free(net)


def chunk_edge_pairs():
    """Pairs whose CIGARs hold runs longer than 2048 moves and runs ending at move 2048 or 4096 exactly: the
    CIGAR text is made 2048 moves at a time, a run carried from one piece to the next."""
    rng = random.Random(57)
    base = "".join(rng.choice("ACGT") for _ in range(6200))
    other = {"A": "C", "C": "G", "G": "T", "T": "A"}
    refs, qrys = [], []
    refs.append(base[:5000]); qrys.append(base[:5000])                                        # one run of 5000
    refs.append(base[:4100]); qrys.append(base[:2048] + other[base[2048]] + base[2049:4100])   # X at 2048
    refs.append(base[:4100]); qrys.append(base[:2047] + other[base[2047]] + base[2048:4100])   # X ends at 2048
    refs.append(base[:4096]); qrys.append(base[:2047] + base[2048:4096])                       # D at 2047
    refs.append(base[:4097]); qrys.append(base[:4095] + other[base[4095]] + base[4096])        # X ends at 4096
    refs.append(base[:3000]); qrys.append(base[:3000] + base[3100:6100])                      # 3000 I
    refs.append(base[:2100]); qrys.append(base[:20])                                           # 2080 D
    refs.append(base[:2048]); qrys.append(base[:2048])                                         # exactly 2048
    return refs, qrys


@pytest.mark.parametrize("right", [False, True])
@pytest.mark.parametrize("eqx", [True, False])
def test_cigar_runs_across_chunk_edges(right, eqx):
    refs, qrys = chunk_edge_pairs()
    b = Batch(refs, qrys, 1, ties_right=right, eqx=eqx)
    try:
        got = b.run()
    finally:
        b.free()
    assert got == oracle_alignments(refs, qrys, right=right, eqx=eqx)
    if eqx:
        assert got[0] == (0, "5000=") and got[1][1] == "2048=1X2051=" and got[7] == (0, "2048=")
    else:
        assert got[0] == (0, "5000M") and got[1][1] == "4100M"

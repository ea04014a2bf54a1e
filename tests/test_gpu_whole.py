"""The whole-matrix path (full.hip) through the C ABI, against the oracle bit for bit.

Every pair of a batch can be sent down it (dinara_batch_set_whole), which checks its recurrence, decisions and walk
on pairs the oracle's whole matrix finishes at once: every cost form (unit costs and their multiples, gap-affine,
deletions priced apart, two-piece), both tie rules, with and without eqx, empty sides, caps, and query lengths that
fill one strip of 512 columns, cross strips or end on a strip's edge. Then the pairs that no band settles take it by
themselves: a length difference past the widest band (the oracle's whole matrix is still small), letters past what
the band kernels stage in LDS (the oracle's proven band, reference tests/test_alignment.mojo:3715-3989 aligns pairs
of that size on the device), and unrelated pairs of 36 kbp whose cost needs a band past 16,384 diagonals (checked by
the bit-parallel distance and by pricing the CIGAR, the oracle's matrix being too large for a test).
"""

import ctypes
import random

import pytest

import dinara_align as da
from dinara_align import _lib
from oracle import oracle as o

pytestmark = pytest.mark.gpu

DISTANCES, ALIGNMENTS = 0, 1


def to_oracle(costs: da.Costs) -> o.Costs:
    if costs.opening2 >= 0 or costs.deletion_opening2 >= 0:
        ins2 = (costs.opening2, costs.extension2) if costs.opening2 >= 0 else (costs.opening, costs.extension)
        if costs.deletion_extension:
            dele = (costs.deletion_opening, costs.deletion_extension)
            del2 = (costs.deletion_opening2, costs.deletion_extension2) if costs.deletion_opening2 >= 0 else dele
        else:
            dele, del2 = (costs.opening, costs.extension), ins2
        return o.Costs(costs.mismatch, dele[0], dele[1], costs.opening, costs.extension, 2, del2[0], del2[1],
                       ins2[0], ins2[1])
    if costs.deletion_extension:
        return o.Costs(costs.mismatch, costs.deletion_opening, costs.deletion_extension, costs.opening,
                       costs.extension)
    return o.Costs.affine(costs.mismatch, costs.opening, costs.extension)


class Batch:
    """A resident batch through the C ABI; `whole` sends every pair down the whole-matrix path."""

    def __init__(self, refs, qrys, kind, costs: da.Costs, whole, right=False, eqx=True, max_cost=-1):
        self.h = _lib.lib()
        self.R, self.Q = _lib.Texts(refs), _lib.Texts(qrys)
        self.costs = costs.as_c()
        self.options = _lib.COptions(-(1 << 60), 1 << 60, max_cost, 1 if eqx else 0, 1 if right else 0, 80 << 20)
        status = ctypes.c_int64(0)
        self.n, self.kind = len(refs), kind
        self.ptr = self.h.dinara_batch_new(0, kind, self.n, *self.R.args(), *self.Q.args(), ctypes.byref(self.costs),
                                           None, ctypes.byref(self.options), 4, ctypes.byref(status))
        assert self.ptr, status.value
        assert self.h.dinara_batch_set_whole(self.ptr, 1 if whole else 0) == 0

    def run(self):
        assert self.h.dinara_batch_run(self.ptr) == 0
        if self.kind == DISTANCES:
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

    def stat(self, what):
        return self.h.dinara_batch_stat(self.ptr, what)

    def free(self):
        self.h.dinara_batch_free(self.ptr)


def run_batch(refs, qrys, kind, costs, whole=True, **kw):
    b = Batch(refs, qrys, kind, costs, whole, **kw)
    try:
        got = b.run()
        return got, b.stat(10)
    finally:
        b.free()


def tie_rich(rng, count, longest):
    refs, qrys = [], []
    for i in range(count):
        alphabet = "AC" if i % 3 == 0 else "ACGT"
        unit = "".join(rng.choice(alphabet) for _ in range(rng.randint(1, 4)))
        n = rng.randint(0, longest)
        base = (unit * (n // len(unit) + 1))[:n] if i % 2 else "".join(rng.choice(alphabet) for _ in range(n))
        q = list(base)
        for _ in range(rng.randint(0, 8)):
            at = rng.randint(0, len(q))
            kind = rng.random()
            if kind < 0.35 and q:
                del q[at:at + rng.randint(1, 30)]
            elif kind < 0.7:
                q[at:at] = list(rng.choice(alphabet) * rng.randint(1, 30))
            elif q:
                q[min(at, len(q) - 1)] = rng.choice(alphabet)
        refs.append(base)
        qrys.append("".join(q))
    # Query lengths on and around a strip's edges, and empty sides.
    for m in (511, 512, 513, 1024, 1025):
        q = "".join(rng.choice("ACGT") for _ in range(m))
        refs.append(q[: m - 7] + "".join(rng.choice("ACGT") for _ in range(9)))
        qrys.append(q)
    refs += ["", "ACGT", "", "A"]
    qrys += ["", "", "ACGTT", "C"]
    return refs, qrys


COSTS = [da.Costs.edit(), da.Costs.linear(2, 2), da.Costs.affine(4, 6, 2), da.Costs.affine(3, 1, 4),
         da.Costs.affine(4, 2, 1).with_deletions(6, 3), da.Costs.two_piece(4, 6, 2, 24, 1),
         da.Costs.two_piece(2, 2, 2, 4, 1), da.Costs.affine(4, 6, 2).with_deletions(8, 2, 20, 1)]


@pytest.mark.parametrize("costs", COSTS, ids=[str(i) for i in range(len(COSTS))])
def test_every_pair_down_the_whole_matrix_against_the_oracle(costs):
    rng = random.Random(hash(costs) & 0xFFFF)
    refs, qrys = tie_rich(rng, 70, 900)
    oc = to_oracle(costs)
    for right in (False, True):
        for eqx in ((True, False) if right else (True,)):
            got, swept = run_batch(refs, qrys, ALIGNMENTS, costs, right=right, eqx=eqx)
            assert swept == len(refs)
            for i, (r, q) in enumerate(zip(refs, qrys)):
                want = o.costs_align(r, q, oc, right_ties=right, eqx=eqx)
                assert got[i] == want, (i, len(r), len(q), right, eqx)
    got, _ = run_batch(refs, qrys, DISTANCES, costs)
    assert got == [o.costs_cost(r, q, oc) for r, q in zip(refs, qrys)]


def test_whole_matrix_caps():
    rng = random.Random(5)
    refs, qrys = tie_rich(rng, 30, 400)
    costs = da.Costs.affine(4, 6, 2)
    oc = to_oracle(costs)
    want = [o.costs_cost(r, q, oc) for r, q in zip(refs, qrys)]
    got, _ = run_batch(refs, qrys, DISTANCES, costs, max_cost=40)
    assert got == [w if w <= 40 else _lib.ABOVE_MAX for w in want]
    got, _ = run_batch(refs, qrys, ALIGNMENTS, costs, max_cost=40)
    for i, (r, q) in enumerate(zip(refs, qrys)):
        if want[i] > 40:
            assert got[i] == _lib.ABOVE_MAX, i
        else:
            assert got[i] == o.costs_align(r, q, oc), i


def test_pairs_no_band_settles():
    rng = random.Random(9)
    # Length differences past 16,384 diagonals, under unit, gap-affine and two-piece costs: whole matrices the oracle
    # fills at once.
    for costs in (da.Costs.edit(), da.Costs.affine(4, 6, 2), da.Costs.two_piece(4, 6, 2, 24, 1)):
        oc = to_oracle(costs)
        refs, qrys = [], []
        for n, m in ((20_000, 600), (600, 18_000)):
            shorter = "".join(rng.choice("ACGT") for _ in range(min(n, m)))
            longer = "".join(rng.choice("ACGT") for _ in range(max(n, m) - min(n, m) - 40))
            cut = rng.randint(0, len(longer))
            full = longer[:cut] + shorter + longer[cut:] + "".join(rng.choice("ACGT") for _ in range(40))
            refs.append(full if n > m else shorter)
            qrys.append(shorter if n > m else full)
        refs.append("ACGTACGTTTGCA")  # a short pair beside them settles in its band as before
        qrys.append("ACGTCGTTTTGCA")
        for right in (False, True):
            got, swept = run_batch(refs, qrys, ALIGNMENTS, costs, whole=False, right=right)
            assert swept == 2, costs
            for i, (r, q) in enumerate(zip(refs, qrys)):
                assert got[i] == o.costs_align(r, q, oc, right_ties=right), (costs, i, right)
        got, swept = run_batch(refs, qrys, DISTANCES, costs, whole=False)
        assert got == [o.costs_cost(r, q, oc) for r, q in zip(refs, qrys)]


def test_pairs_past_the_band_kernels_letters():
    # 68,000 and 140,000 letters a pair (a 70 kbp pair, as the reference's device tests align), past the 64 KiB the
    # gap-affine band kernels stage: close pairs, whose proven band the oracle traces in little memory.
    rng = random.Random(10)
    costs = da.Costs.affine(4, 6, 2)
    oc = to_oracle(costs)
    refs, qrys = [], []
    for length, edits in ((34_000, 300), (70_000, 250)):
        r = "".join(rng.choice("ACGT") for _ in range(length))
        q = list(r)
        for _ in range(edits):
            at = rng.randint(0, len(q) - 1)
            kind = rng.random()
            if kind < 0.4:
                q[at] = rng.choice("ACGT")
            elif kind < 0.7:
                del q[at:at + rng.randint(1, 5)]
            else:
                q[at:at] = [rng.choice("ACGT") for _ in range(rng.randint(1, 5))]
        refs.append(r)
        qrys.append("".join(q))
    got, swept = run_batch(refs, qrys, ALIGNMENTS, costs, whole=False)
    assert swept == 2
    for i, (r, q) in enumerate(zip(refs, qrys)):
        assert got[i] == o.costs_align_proven(r, q, oc), i
    got, _ = run_batch(refs, qrys, DISTANCES, costs, whole=False)
    assert got == [o.costs_distance_proven(r, q, oc) for r, q in zip(refs, qrys)]


def test_unrelated_long_pairs_past_the_widest_bin():
    # Two unrelated 36 kbp sequences cost about 19,000 at unit costs: their band passes the widest bin (16,384).
    rng = random.Random(11)
    refs = ["".join(rng.choice("ACGT") for _ in range(36_000))]
    qrys = ["".join(rng.choice("ACGT") for _ in range(36_100))]
    want = o.myers_distance(refs[0], qrys[0])
    assert want > 16_384
    got, swept = run_batch(refs, qrys, DISTANCES, da.Costs.edit(), whole=False)
    assert (got, swept) == ([want], 1)
    got, swept = run_batch(refs, qrys, ALIGNMENTS, da.Costs.edit(), whole=False)
    assert swept == 1 and got[0][0] == want
    assert o.cigar_cost(refs[0], qrys[0], o.Costs.affine(1, 0, 1), got[0][1]) == want

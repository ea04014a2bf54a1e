"""GPU parity under Costs: distances and alignments from libdinara_hip.so (through the C ABI) against the oracle,
pair by pair, bit for bit: the cost and the CIGAR the reference's tie rule spells.

Cases follow the reference's device tests (tests/test_alignment.mojo:3715-3989: the device against the host pair by
pair) and its known answers, the benchmark configurations' shapes, and the edges: empty and ragged pairs,
letters outside ACGT, bands, caps, both tie rules, `M` CIGARs, and bands wide enough for the wave kernel.
"""

import ctypes
import json
import os
import random

import pytest

import dinara_align as da
from dinara_align import synthetic
from oracle import oracle as o

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def to_oracle(costs: da.Costs) -> o.Costs:
    if costs.deletion_extension:
        return o.Costs(costs.mismatch, costs.deletion_opening, costs.deletion_extension, costs.opening,
                       costs.extension)
    return o.Costs.affine(costs.mismatch, costs.opening, costs.extension)


def mutated(text: str, rate: float, rng: random.Random, alphabet: str = "ACGT") -> str:
    out = []
    for ch in text:
        u = rng.random()
        if u < rate / 3:
            out.append(rng.choice([c for c in alphabet if c != ch] or [ch]))
        elif u < 2 * rate / 3:
            out.append(rng.choice(alphabet))
            out.append(ch)
        elif u < rate:
            continue
        else:
            out.append(ch)
    return "".join(out)


def random_pairs(count, rng, lengths=(0, 80), rates=(0.0, 0.05, 0.2, 0.5), alphabet="ACGT"):
    refs, qrys = [], []
    for i in range(count):
        r = "".join(rng.choice(alphabet) for _ in range(rng.randint(*lengths)))
        refs.append(r)
        qrys.append(mutated(r, rates[i % len(rates)], rng, alphabet))
    return refs, qrys


def check_alignments(refs, qrys, costs, ties="left", eqx=True, band=None, max_cost=None):
    got = da.alignments(refs, qrys, costs, ties=ties, eqx=eqx, band=band, max_cost=max_cost)
    oc = to_oracle(costs)
    oband = (band.low, band.high) if band else (o.NO_BAND_LOW, o.NO_BAND_HIGH)
    for i, (r, q) in enumerate(zip(refs, qrys)):
        cost, cigar = o.costs_align(r, q, oc, oband, right_ties=ties == "right", eqx=eqx)
        if max_cost is not None and (cost > max_cost or cost == o.OUTSIDE):
            assert got[i] is None, (i, r, q)
            continue
        assert got[i] is not None, (i, r, q)
        assert (got[i].cost, got[i].cigar) == (cost, cigar), (i, r, q, costs, ties)
        assert (got[i].reference_start, got[i].reference_end, got[i].query_start, got[i].query_end) == (
            0, len(r), 0, len(q))
        assert got[i].score == -cost


def check_distances(refs, qrys, costs, band=None, max_cost=None):
    got = da.distances(refs, qrys, costs, band=band, max_cost=max_cost)
    oc = to_oracle(costs)
    oband = (band.low, band.high) if band else (o.NO_BAND_LOW, o.NO_BAND_HIGH)
    for i, (r, q) in enumerate(zip(refs, qrys)):
        cost = o.costs_cost(r, q, oc, oband)
        # Under a cap, a band no alignment fits leaves nothing within it (reference c/dinara.mojo:612-615).
        want = None if (max_cost is not None and (cost > max_cost or cost == o.OUTSIDE)) else cost
        assert got[i] == want, (i, r, q, costs)


def test_known_answers():
    # tests/test_python.py:70-79, tests/test_c_api.cpp:69-158
    assert da.distance("ACGTACGTTTGCA", "ACGTCGTTTTGCA") == 2
    found = da.align("ACGTACGTTTGCA", "ACGTCGTTTTGCA")
    assert (found.cost, found.cigar) == (2, "4=1D2=1I6=")
    assert da.align("ACGTACGTTTGCA", "ACGTCGTTTTGCA", eqx=False).cigar == "4M1D2M1I6M"
    assert da.align("ACGTACGTTTGCA", "ACGTCGTTTTGCA", ties="right").cigar == "4=1D5=1I3="
    affine = da.Costs.affine(4, 6, 2)
    assert da.align("ACGTTTTACG", "ACGTTTACG", affine).cigar == "3=1D6="
    assert da.align("ACGTTTTACG", "ACGTTTACG", affine, ties="right").cigar == "6=1D3="
    assert da.align(b"ACGT", b"ACGT").cigar == "4="
    assert da.distance("ACGTACGTTTGCA", "ACGTCGTTTTGCA", affine, max_cost=11) is None
    assert da.align("ACGTACGTTTGCA", "ACGTCGTTTTGCA", affine, max_cost=12).cigar == "4=3X6="
    assert da.distance("AAAACCCC", "CCCCAAAA", affine, band=da.Band.around(0)) == 32
    assert da.distance("AAAACCCC", "CCCCAAAA", da.Costs.edit(), band=da.Band.around(0)) == 8
    assert da.align("ACGTACGT", "ACGAACGT", affine, band=da.Band(0, 0)).cigar == "3=1X4="
    assert da.score("ACGTACGTTTGCA", "ACGTCGTTTTGCA", affine) == -12
    assert da.distance("ACGACGT", "ACGTACGT", affine.with_deletions(20, 5)) == 8
    assert da.distance("ACGTACGT", "ACGACGT", affine.with_deletions(20, 5)) == 25
    dear = da.Costs.affine(4, 2, 1).with_deletions(20, 5)
    assert da.align("ACGACGT", "ACGTACGT", dear).cigar == "3=1I4="
    assert da.align("", "ACG").cigar == "3I"
    assert da.align("ACG", "").cigar == "3D"
    empty = da.align("", "")
    assert empty.cost == 0 and empty.cigar == ""
    assert da.align("ACGT", "", affine).cost == 14
    assert da.align("ACGTNNAC", "ACGTNNAC").cigar == "8="
    assert da.distance("NRYKMSW", "NRYKMSA") == 1


def test_refusals():
    with pytest.raises(ValueError):
        da.align("A", "C", da.Costs.affine(0, 6, 2))
    with pytest.raises(ValueError):
        da.distance("ACGT", "AC", da.Costs.affine(4, 6, 2), band=da.Band.around(1))
    assert da.distance("ACGT", "AC", da.Costs.affine(4, 6, 2), band=da.Band.around(1), max_cost=100) is None
    with pytest.raises(ValueError):
        da.distance("\xfe".encode("latin-1"), b"A")
    with pytest.raises(ValueError):
        da.distance("ACGT", "ACGT", da.Costs.affine(4, 6, 2), da.Mode.local(2))
    assert da.distance("ACGTACGTTTGCA", "ACGTCGTTTTGCA", max_cost=-1) is None


COSTS = [da.Costs.edit(), da.Costs.affine(4, 6, 2), da.Costs.affine(4, 2, 1).with_deletions(9, 1),
         da.Costs.linear(3, 2), da.Costs.affine(1, 0, 1), da.Costs.affine(5, 11, 1)]


@pytest.mark.parametrize("costs", COSTS)
def test_random_pairs_alignments(costs):
    rng = random.Random(hash((costs.mismatch, costs.opening, costs.extension)) & 0xFFFF)
    refs, qrys = random_pairs(300, rng)
    check_alignments(refs, qrys, costs)
    check_alignments(refs[:100], qrys[:100], costs, ties="right")
    check_alignments(refs[:60], qrys[:60], costs, eqx=False)


@pytest.mark.parametrize("costs", COSTS[:3])
def test_random_pairs_distances(costs):
    rng = random.Random(7)
    refs, qrys = random_pairs(400, rng, lengths=(0, 200))
    check_distances(refs, qrys, costs)


def test_letters_outside_acgt():
    rng = random.Random(21)
    # At most four distinct bytes a pair: two bits a letter, each pair its own codes.
    refs, qrys = random_pairs(100, rng, alphabet="ACGN")
    check_alignments(refs, qrys, da.Costs.affine(4, 6, 2))
    # Twenty letters: a byte a letter.
    refs, qrys = random_pairs(100, rng, alphabet="ACDEFGHIKLMNPQRSTVWY")
    check_alignments(refs, qrys, da.Costs.affine(4, 6, 2))
    check_distances(refs, qrys, da.Costs.edit())


def test_ragged_lengths_and_empties():
    rng = random.Random(23)
    refs, qrys = [], []
    for _ in range(150):
        r = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 300)))
        q = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 300)))
        refs.append(r)
        qrys.append(q)
    refs += ["", "A", "", "ACGT" * 20]
    qrys += ["", "", "A", ""]
    check_alignments(refs, qrys, da.Costs.affine(4, 6, 2))
    check_alignments(refs, qrys, da.Costs.edit(), ties="right")


def test_wide_bands():
    rng = random.Random(29)
    refs, qrys = random_pairs(40, rng, lengths=(300, 900), rates=(0.3, 0.5, 0.8))
    # Length differences the lanes' widest band cannot hold.
    refs += ["ACGT" * 100, "A" * 700]
    qrys += ["ACGT" * 10, "C" * 150]
    check_alignments(refs, qrys, da.Costs.affine(4, 6, 2))
    check_alignments(refs, qrys, da.Costs.edit(), ties="right")
    check_distances(refs, qrys, da.Costs.edit())


def class_pairs(refs, qrys, costs, kind):
    """Runs a resident batch through the C ABI and returns how many pairs each band class swept."""
    from dinara_align import _lib
    handle = _lib.lib()
    R, Q = _lib.Texts(refs), _lib.Texts(qrys)
    c = costs.as_c()
    status = ctypes.c_int64(0)
    batch = handle.dinara_batch_new(0, kind, len(refs), *R.args(), *Q.args(), ctypes.byref(c), None, None, 4,
                                    ctypes.byref(status))
    assert batch, status.value
    try:
        assert handle.dinara_batch_run(batch) == 0
        return [int(handle.dinara_batch_stat(batch, 400 + cls)) for cls in range(22)]
    finally:
        handle.dinara_batch_free(batch)


def resident_alignments(refs, qrys, costs):
    """A resident batch's alignments through the C ABI, and stat 14 (above zero: packed 16-bit group cells)."""
    from dinara_align import _lib
    handle = _lib.lib()
    R, Q = _lib.Texts(refs), _lib.Texts(qrys)
    c = costs.as_c()
    status = ctypes.c_int64(0)
    batch = handle.dinara_batch_new(0, _lib.BATCH_ALIGNMENTS, len(refs), *R.args(), *Q.args(), ctypes.byref(c), None,
                                    None, 4, ctypes.byref(status))
    assert batch, status.value
    try:
        assert handle.dinara_batch_run(batch) == 0
        found = (_lib.CAlignment * len(refs))()
        statuses = (ctypes.c_int64 * len(refs))()
        assert handle.dinara_batch_alignments(batch, found, statuses) == 0
        out = []
        for i in range(len(refs)):
            assert statuses[i] == 0, (i, statuses[i])
            out.append((found[i].cost, ctypes.string_at(found[i].cigar, found[i].cigar_length).decode()))
            handle.dinara_free(found[i].cigar)
        return out, handle.dinara_batch_stat(batch, 14)
    finally:
        handle.dinara_batch_free(batch)


@pytest.mark.parametrize("costs", [da.Costs.affine(4, 6, 2), da.Costs.affine(2, 3, 1), da.Costs.affine(1, 5, 3)])
def test_packed_group_cells(costs, monkeypatch):
    # Shifted costs whose mismatch is at most two extension letters: the group sweeps hold two slots a register as
    # 16-bit halves while the batch's longest group pair stays under a bound (about 1,000 letters at (4, 6, 2)).
    # Pairs up to that bound, tie-rich repeats, unrelated and ragged pairs, against the oracle and the 32-bit cells;
    # a batch past the bound takes the 32-bit cells.
    rng = random.Random(costs.mismatch * 100 + costs.opening * 10 + costs.extension)
    limit = {(4, 6, 2): 1000, (2, 3, 1): 2000, (1, 5, 3): 640}[(costs.mismatch, costs.opening, costs.extension)]
    refs, qrys = random_pairs(60, rng, lengths=(limit // 2 - 60, limit // 2 - 30), rates=(0.01, 0.05, 0.15, 0.3))
    for unit in ("A", "AC", "ACG", "AACC"):
        r = (unit * limit)[: limit // 2 - 20]
        refs.append(r)
        qrys.append(mutated(r, 0.1, rng))
    for _ in range(8):
        refs.append("".join(rng.choice("ACGT") for _ in range(rng.randint(limit // 2 - 100, limit // 2 - 20))))
        qrys.append("".join(rng.choice("ACGT") for _ in range(rng.randint(limit // 2 - 100, limit // 2 - 20))))
    refs += ["ACGT" * 3, "", "A" * (limit // 2 - 10)]
    qrys += ["AGT", "ACG", "A" * (limit // 2 - 40)]
    assert max(len(r) + len(q) for r, q in zip(refs, qrys)) <= limit
    oc = to_oracle(costs)
    want = [o.costs_align(r, q, oc) for r, q in zip(refs, qrys)]
    packed, top = resident_alignments(refs, qrys, costs)
    assert top > 0
    assert packed == want
    monkeypatch.setenv("DINARA_GROUP_CELLS32", "1")
    wide, top = resident_alignments(refs, qrys, costs)
    assert top == 0 and wide == want
    monkeypatch.delenv("DINARA_GROUP_CELLS32")
    # Past the bound: one pair of 60 letters more than it.
    refs.append("ACGT" * ((limit + 60) // 8))
    qrys.append(mutated(refs[-1], 0.05, rng))
    beyond, top = resident_alignments(refs, qrys, costs)
    assert top == 0
    assert beyond[:-1] == want and beyond[-1] == o.costs_align(refs[-1], qrys[-1], oc)


def test_block_bands_one_and_several_waves():
    # Long, noisy pairs: bands of 512 to 8192 diagonals, one pair a block of one to eight waves.
    rng = random.Random(37)
    refs, qrys = [], []
    for n, rate in [(1200, 0.25), (2500, 0.3), (3000, 0.45), (1800, 0.6), (4000, 0.5), (900, 0.8)]:
        r = "".join(rng.choice("ACGT") for _ in range(n))
        refs.append(r)
        qrys.append(mutated(r, rate, rng))
    # Length differences that put the first band past one wave's.
    refs += ["ACGT" * 700, "G" * 2500, "ACG" * 900]
    qrys += ["ACGT" * 40, "GT" * 200, "TTAC" * 300]
    for costs in (da.Costs.affine(4, 6, 2), da.Costs.linear(3, 2)):
        swept = class_pairs(refs, qrys, costs, 1)
        assert sum(swept[10:22]) > 0 and sum(swept[12:22]) > 0, swept  # blocks of several waves ran
        check_alignments(refs, qrys, costs)
        check_distances(refs, qrys, costs)
    # Unit costs take the bit-parallel path (test_gpu_edit.py); the same pairs there.
    check_alignments(refs, qrys, da.Costs.edit())
    check_distances(refs, qrys, da.Costs.edit())


def test_widest_band_distances():
    # First bands past 8192 diagonals: the 12288- and 16384-diagonal blocks, distances only.
    rng = random.Random(41)
    refs = ["".join(rng.choice("ACGT") for _ in range(n)) for n in (11000, 9500)]
    qrys = [refs[0][:1500], mutated(refs[1][2000:3200], 0.1, rng)]
    swept = class_pairs(refs, qrys, da.Costs.affine(4, 6, 2), 0)
    assert swept[20] + swept[21] > 0, swept
    check_distances(refs, qrys, da.Costs.affine(4, 6, 2))
    check_distances(refs, qrys, da.Costs.edit())


def test_user_bands_and_caps():
    rng = random.Random(31)
    refs, qrys = random_pairs(120, rng, lengths=(1, 120), rates=(0.05, 0.2))
    for band in (da.Band.around(3), da.Band(-10, 2), da.Band(0, 40)):
        check_distances(refs, qrys, da.Costs.affine(4, 6, 2), band=band, max_cost=10_000)
    check_alignments(refs, qrys, da.Costs.affine(4, 6, 2), band=da.Band.around(8), max_cost=10_000)
    check_alignments(refs, qrys, da.Costs.affine(4, 6, 2), max_cost=30)
    check_distances(refs, qrys, da.Costs.edit(), max_cost=5)


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_golden_c1():
    data = golden("c1_edit.json")["pairs"]
    got = da.distances([p[0] for p in data], [p[1] for p in data])
    assert got == [p[2] for p in data]


def test_golden_c2():
    data = golden("c2_affine.json")["pairs"]
    affine = da.Costs.affine(4, 6, 2)
    got = da.alignments([p[0] for p in data], [p[1] for p in data], affine)
    assert [(a.cost, a.cigar) for a in got] == [(p[2], p[3]) for p in data]
    right = [p for p in data if len(p) > 4]
    got = da.alignments([p[0] for p in right], [p[1] for p in right], affine, ties="right")
    assert [a.cigar for a in got] == [p[4] for p in right]


def test_golden_c3_c4():
    data = golden("c3_uneven_small.json")["pairs"]
    got = da.alignments([p[0] for p in data], [p[1] for p in data])
    assert [(a.cost, a.cigar) for a in got] == [(p[2], p[3]) for p in data]
    data = golden("c4_affine_uneven_small.json")["pairs"]
    got = da.distances([p[0] for p in data], [p[1] for p in data], da.Costs.affine(4, 6, 2))
    assert got == [p[2] for p in data]


def rows_cost(cigar, ref, qry, costs):
    """What a CIGAR's alignment costs, each gap run at its opening and extensions: a rescorer of its own."""
    total, i, j, num = 0, 0, 0, ""
    for ch in cigar:
        if ch.isdigit():
            num += ch
            continue
        k = int(num)
        num = ""
        if ch == "D":
            total += costs.opening + k * costs.extension
            i += k
        elif ch == "I":
            total += costs.opening + k * costs.extension
            j += k
        else:
            for _ in range(k):
                same = ref[i] == qry[j]
                assert same == (ch == "=")
                total += 0 if same else costs.mismatch
                i += 1
                j += 1
    assert i == len(ref) and j == len(qry)
    return total


def test_full_size_c2_properties():
    """Configuration 2 at full size: every CIGAR rebuilds both sequences and prices at the reported cost, each
    cost matches the distance call, and a seeded sample matches the oracle bit for bit."""
    refs, qrys = synthetic.affine_reads(100_000)
    affine = da.Costs.affine(4, 6, 2)
    found = da.alignments(refs, qrys, affine, threads=8)
    costs = da.distances(refs, qrys, affine, threads=8)
    assert [a.cost for a in found] == costs
    for i in range(0, 100_000, 97):
        assert rows_cost(found[i].cigar, refs[i], qrys[i], affine) == found[i].cost
    oc = o.Costs.affine(4, 6, 2)
    for i in range(0, 100_000, 1009):
        assert (found[i].cost, found[i].cigar) == o.costs_align(refs[i], qrys[i], oc)


def test_c_abi_codes_per_pair():
    """Per-pair codes straight through the C ABI (reference c/dinara.h:35-49, c/dinara.mojo:96-108)."""
    import ctypes
    from dinara_align import _lib
    handle = _lib.lib()
    refs = _lib.Texts([b"ACGT", b"AC\xfeT", b"ACGT", b"ACGTACGT"])
    qrys = _lib.Texts([b"ACGA", b"ACGT", b"ACGT", b"ACGT"])
    qrys.lengths[2] = -1
    results = (ctypes.c_int64 * 4)()
    c = _lib.CCosts(4, 6, 2, -1, 0, 0, 0, -1, 0)
    assert handle.dinara_distances(4, *refs.args(), *qrys.args(), ctypes.byref(c), None, None, 0, results) == 0
    assert list(results) == [4, _lib.UNSUPPORTED_SYMBOLS, _lib.INVALID_LENGTH, 14]
    bad = _lib.CCosts(4, -1, 2, -1, 0, 0, 0, -1, 0)
    results[0] = 12345
    assert handle.dinara_distances(4, *refs.args(), *qrys.args(), ctypes.byref(bad), None, None, 0,
                                   results) == _lib.INVALID_COSTS
    assert results[0] == 12345  # nothing written on a batch-wide failure
    assert handle.dinara_distances(0, *refs.args(), *qrys.args(), None, None, None, 0, results) == 0
    found = (_lib.CAlignment * 4)()
    statuses = (ctypes.c_int64 * 4)()
    assert handle.dinara_alignments(4, *refs.args(), *qrys.args(), ctypes.byref(c), None, None, 2, found,
                                    statuses) == 0
    assert list(statuses) == [0, _lib.UNSUPPORTED_SYMBOLS, _lib.INVALID_LENGTH, 0]
    assert ctypes.string_at(found[0].cigar).decode() == "3=1X"
    assert ctypes.string_at(found[3].cigar).decode() == "4D4="
    assert found[1].cigar is None
    for item in found:
        handle.dinara_free(item.cigar)


def test_cigar_runs_across_chunk_edges():
    # Runs longer than 2048 moves and runs ending at move 2048 exactly: the CIGAR text is made 2048 moves at a time.
    rng = random.Random(59)
    base = "".join(rng.choice("ACGT") for _ in range(5200))
    other = {"A": "C", "C": "G", "G": "T", "T": "A"}
    refs = [base[:5000], base[:4100], base[:4096], base[:3000]]
    qrys = [base[:5000], base[:2047] + other[base[2047]] + base[2048:4100], base[:2047] + base[2048:4096],
            base[:3000] + base[4000:5100]]
    for costs in (da.Costs.affine(4, 6, 2), da.Costs.linear(3, 2)):
        check_alignments(refs, qrys, costs)
        check_alignments(refs, qrys, costs, ties="right", eqx=False)

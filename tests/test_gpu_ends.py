"""Free ends on the device (the wave sweep's forward phase for the cost and end, its anchored phase for the start
and the walk), through the package and the C ABI, against the oracle's free-end rule pair by pair, bit for bit:
cost, CIGAR and the four span coordinates.

Known answers are the reference's (tests/test_c_api.cpp:111-121, tests/test_cli.sh:28-29); random and tie-rich
pairs follow its check of the span rule (tests/test_alignment.mojo:2023-2113): repeats, free letters of 0, a few,
and every letter, under one-piece, two-piece and unit costs, both tie rules.
"""

import random

import pytest

import dinara_align as da
from dinara_align.types import UNBOUNDED
from oracle import oracle as o

pytestmark = pytest.mark.gpu

AFFINE = da.Costs.affine(4, 6, 2)


def to_oracle(costs: da.Costs) -> o.Costs:
    if costs.opening2 >= 0:
        return o.Costs.two_piece(costs.mismatch, costs.opening, costs.extension, costs.opening2, costs.extension2)
    if costs.deletion_extension:
        return o.Costs(costs.mismatch, costs.deletion_opening, costs.deletion_extension, costs.opening,
                       costs.extension)
    return o.Costs.affine(costs.mismatch, costs.opening, costs.extension)


def ends_of(mode: da.Mode):
    return (mode.reference_start, mode.reference_end, mode.query_start, mode.query_end)


def test_known_answers():
    placed = da.align("TTTTACGTACGTTTTT", "ACGTACGT", AFFINE, da.Mode.INFIX)
    assert (placed.cost, placed.cigar, placed.reference_start, placed.reference_end) == (0, "8=", 4, 12)
    found = da.align("TTTTACGTACGTTTTT", "ACGTCGT", da.Costs.edit(), da.Mode.INFIX)
    assert found.cost == 1 and found.reference_start == 4
    assert da.distance("TTTTACGTACGTTTTT", "ACGTACGT", AFFINE) > 0
    assert da.distance("TTTTACGTACGTTTTT", "ACGTACGT", AFFINE, da.Mode.INFIX) == 0
    assert da.distance("ACGTACGTTTTT", "ACGTACGT", AFFINE, da.Mode.PREFIX) == 0
    assert da.distance("TTTTACGTACGT", "ACGTACGT", AFFINE, da.Mode.SUFFIX) == 0
    assert da.distance("TTTTACGTACGT", "ACGTACGT", da.Costs.edit(), da.Mode.SUFFIX) == 0
    overlap = da.align("TTTTTACGTACGT", "ACGTACGTGGGGG", AFFINE, da.Mode.ends_free(5, 0, 0, 5))
    assert (overlap.cost, overlap.cigar, overlap.reference_start, overlap.query_end) == (0, "8=", 5, 8)
    chrom = "TTTTTTTTTTACGTACGTTTGCAGGATCCAGTTTTTTTTTT"
    read = da.align(chrom, "ACGTACGTTTGCAGGATCCAG", AFFINE, da.Mode.INFIX)
    assert (read.cost, read.reference_start, read.reference_end, read.query_start, read.query_end, read.cigar) == (
        0, 10, 31, 0, 21, "21=")
    with pytest.raises(ValueError):  # free ends under a user band are not served
        da.align(chrom, "ACGT", AFFINE, da.Mode.INFIX, band=da.Band(-50, 50))


def tie_rich(rng, count):
    refs, qrys = [], []
    for i in range(count):
        alphabet = "AC" if i % 3 == 0 else "ACGT"
        unit = "".join(rng.choice(alphabet) for _ in range(rng.randint(1, 4)))
        core = "".join(rng.choice(alphabet) for _ in range(rng.randint(0, 60)))
        for _ in range(rng.randint(0, 5)):
            core += unit * rng.randint(1, 4) + "".join(rng.choice(alphabet) for _ in range(rng.randint(0, 15)))
        q = list(core[rng.randint(0, len(core) // 3):len(core) - rng.randint(0, len(core) // 3)])
        for _ in range(rng.randint(0, 6)):
            at = rng.randint(0, len(q))
            kind = rng.random()
            if kind < 0.35 and q:
                del q[at:at + rng.randint(1, 4)]
            elif kind < 0.7:
                q[at:at] = list(rng.choice(alphabet) * rng.randint(1, 4))
            elif q:
                q[min(at, len(q) - 1)] = rng.choice(alphabet)
        if rng.random() < 0.3:
            q = list("".join(rng.choice(alphabet) for _ in range(rng.randint(0, 12)))) + q
        refs.append(core)
        qrys.append("".join(q))
    return refs, qrys


COSTS = [AFFINE, da.Costs.edit(), da.Costs.two_piece(4, 6, 2, 24, 1), da.Costs.affine(3, 1, 4),
         da.Costs.affine(4, 2, 1).with_deletions(6, 3)]
MODES = [da.Mode.INFIX, da.Mode.PREFIX, da.Mode.SUFFIX, da.Mode.ends_free(2, 9, 0, 3),
         da.Mode.ends_free(0, 0, UNBOUNDED, UNBOUNDED), da.Mode.ends_free(5, 0, 0, 5),
         da.Mode.ends_free(1, 1, 1, 1)]


@pytest.mark.parametrize("costs", COSTS, ids=["affine", "edit", "two-piece", "linear-ish", "deletions"])
def test_random_free_ends_against_the_oracle(costs):
    rng = random.Random(hash(costs) & 0xFFFF)
    oc = to_oracle(costs)
    for mode in MODES:
        refs, qrys = tie_rich(rng, 60)
        ends = ends_of(mode)
        for ties in ("left", "right"):
            got = da.alignments(refs, qrys, costs, mode, ties=ties)
            for i, (r, q) in enumerate(zip(refs, qrys)):
                cost, cigar, spans = o.ends_align(r, q, oc, ends, right_ties=ties == "right")
                g = got[i]
                assert g is not None, (i, r, q, mode, ties)
                assert (g.cost, g.cigar, (g.reference_start, g.reference_end, g.query_start, g.query_end)) == (
                    cost, cigar, spans), (i, r, q, mode, ties)
        distances = da.distances(refs, qrys, costs, mode)
        assert distances == [o.ends_cost(r, q, oc, ends)[0] for r, q in zip(refs, qrys)], mode


def test_free_ends_caps_empty_and_longer_pairs():
    rng = random.Random(17)
    oc = to_oracle(AFFINE)
    refs, qrys = [], []
    for _ in range(40):
        window = "".join(rng.choice("ACGT") for _ in range(rng.randint(300, 900)))
        at = rng.randint(0, len(window) - 150)
        read = list(window[at:at + rng.randint(80, 150)])
        for _ in range(rng.randint(0, 6)):
            read[rng.randint(0, len(read) - 1)] = rng.choice("ACGT")
        refs.append(window)
        qrys.append("".join(read))
    refs += ["", "ACGT", "", "ACGTACGT"]
    qrys += ["", "", "ACG", "CGTA"]
    got = da.alignments(refs, qrys, AFFINE, da.Mode.INFIX, eqx=False)
    for i, (r, q) in enumerate(zip(refs, qrys)):
        cost, cigar, spans = o.ends_align(r, q, oc, ends_of(da.Mode.INFIX), eqx=False)
        assert (got[i].cost, got[i].cigar, (got[i].reference_start, got[i].reference_end, got[i].query_start,
                                            got[i].query_end)) == (cost, cigar, spans), i
    capped = da.distances(refs, qrys, AFFINE, da.Mode.INFIX, max_cost=10)
    for i, (r, q) in enumerate(zip(refs, qrys)):
        want = o.ends_cost(r, q, oc, ends_of(da.Mode.INFIX))[0]
        assert capped[i] == (want if want <= 10 else None), i
    capped = da.alignments(refs, qrys, AFFINE, da.Mode.INFIX, max_cost=10)
    for i, (r, q) in enumerate(zip(refs, qrys)):
        want = o.ends_align(r, q, oc, ends_of(da.Mode.INFIX))
        assert (capped[i] is None) == (want[0] > 10), i
        if capped[i] is not None:
            assert capped[i].cigar == want[1]

"""Two-piece gap costs on the device (the wave sweep's second deletion and insertion layers, the two-piece
certificate and walk), through the package and the C ABI, against the oracle pair by pair, bit for bit.

Known answers are the reference's (tests/test_c_api.cpp:121-129, tests/test_alignment.mojo:1765-1771); random and
tie-rich pairs follow its full-matrix check of two-piece costs (tests/test_alignment.mojo:1700-1745) under both
tie rules, with deletions priced apart on either side, as reference modes.mojo:99-143 defines them.
"""

import random

import pytest

import dinara_align as da
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def to_oracle(costs: da.Costs) -> o.Costs:
    """The oracle's lanes terms of a two-piece Costs (a side given one piece counts it twice)."""
    ins2 = (costs.opening2, costs.extension2) if costs.opening2 >= 0 else (costs.opening, costs.extension)
    if costs.deletion_extension:
        dele = (costs.deletion_opening, costs.deletion_extension)
        del2 = (costs.deletion_opening2, costs.deletion_extension2) if costs.deletion_opening2 >= 0 else dele
        if costs.opening2 < 0:
            ins2 = (costs.opening, costs.extension)
    else:
        dele, del2 = (costs.opening, costs.extension), ins2
    return o.Costs(costs.mismatch, dele[0], dele[1], costs.opening, costs.extension, 2, del2[0], del2[1], ins2[0],
                   ins2[1])


def test_known_answers():
    gapped = "GATTACAGCTTGCA" + "C" * 30 + "TGGACCATGAGTCATTGACCAGTCGATC"
    plain = "GATTACAGCTTGCATGGACCATGAGTCAGTTGACCAGTCGATC"
    two = da.Costs.two_piece(4, 6, 2, 24, 1)
    found = da.align(gapped, plain, two)
    assert (found.cost, found.cigar) == (62, "14=30D14=1I14=")
    assert da.distance(gapped, plain, two) == 62
    assert da.distance(gapped, plain, da.Costs.affine(4, 6, 2)) == 74
    assert da.distance("", "ACGT", two) == 14
    assert da.distance("", "A" * 30, two) == 54
    assert da.distance(gapped, plain, two, max_cost=61) is None
    assert da.align(gapped, plain, two, max_cost=62).cost == 62


COSTS = [
    da.Costs.two_piece(4, 6, 2, 24, 1),
    da.Costs.two_piece(3, 4, 2, 9, 1),
    da.Costs.two_piece(2, 2, 2, 4, 1),  # the pieces cost alike at two and three letters
    da.Costs.two_piece(3, 1, 2, 4, 1),
    da.Costs.affine(4, 6, 2).with_deletions(8, 2, 20, 1),  # deletions alone have a second piece
    da.Costs.two_piece(4, 4, 2, 24, 1).with_deletions(6, 3),  # insertions alone have one
]


def tie_rich_pairs(count, rng):
    refs, qrys = [], []
    for i in range(count):
        alphabet = "AC" if i % 3 == 0 else ("ACG" if i % 3 == 1 else "ACGT")
        unit = "".join(rng.choice(alphabet) for _ in range(rng.randint(1, 4)))
        base = (unit * 40)[: rng.randint(0, 90)] if i % 2 else "".join(rng.choice(alphabet) for _ in range(
            rng.randint(0, 120)))
        q = list(base)
        for _ in range(rng.randint(0, 6)):
            kind = rng.random()
            at = rng.randint(0, len(q))
            if kind < 0.4 and q:
                del q[at:at + rng.randint(1, 25)]
            elif kind < 0.8:
                q[at:at] = list(rng.choice(alphabet) * rng.randint(1, 25))
            elif q:
                q[min(at, len(q) - 1)] = rng.choice(alphabet)
        refs.append(base)
        qrys.append("".join(q))
    return refs, qrys


@pytest.mark.parametrize("costs", COSTS, ids=[f"{c}" for c in range(len(COSTS))])
def test_random_two_piece_alignments_both_tie_rules(costs):
    rng = random.Random(hash(costs) & 0xFFFF)
    refs, qrys = tie_rich_pairs(300, rng)
    oc = to_oracle(costs)
    for ties in ("left", "right"):
        got = da.alignments(refs, qrys, costs, ties=ties)
        for i, (r, q) in enumerate(zip(refs, qrys)):
            want = o.costs_align(r, q, oc, right_ties=ties == "right")
            assert (got[i].cost, got[i].cigar) == want, (i, r, q, ties)
    distances = da.distances(refs, qrys, costs)
    assert distances == [o.costs_cost(r, q, oc) for r, q in zip(refs, qrys)]


def test_two_piece_long_gaps_and_bands():
    rng = random.Random(7)
    two = da.Costs.two_piece(4, 6, 2, 24, 1)
    oc = to_oracle(two)
    refs, qrys = [], []
    for _ in range(60):
        core = "".join(rng.choice("ACGT") for _ in range(rng.randint(200, 700)))
        cut = rng.randint(0, len(core))
        gap = rng.randint(20, 300)
        if rng.random() < 0.5:
            refs.append(core[:cut] + "".join(rng.choice("ACGT") for _ in range(gap)) + core[cut:])
            qrys.append(core)
        else:
            refs.append(core)
            qrys.append(core[:cut] + "".join(rng.choice("ACGT") for _ in range(gap)) + core[cut:])
    got = da.alignments(refs, qrys, two, eqx=False)
    for i, (r, q) in enumerate(zip(refs, qrys)):
        assert (got[i].cost, got[i].cigar) == o.costs_align_proven(r, q, oc, eqx=False), i
    band = da.Band(-40, 40)
    got = da.distances(refs[:20], qrys[:20], two, band=band, max_cost=10_000)
    for i, (r, q) in enumerate(zip(refs[:20], qrys[:20])):
        want = o.costs_cost(r, q, oc, (band.low, band.high))
        assert got[i] == (None if want == o.OUTSIDE else want), i

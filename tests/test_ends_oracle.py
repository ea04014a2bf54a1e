"""The oracle's free ends (oracle.c orc_ends_cost / orc_ends_align), pinned three ways, no GPU:
- the reference's own answers for infix, prefix, suffix and overlap placements (tests/test_c_api.cpp:111-121,
  tests/test_cli.sh:28-29);
- brute force: every alignment of every allowed start and end of small pairs enumerated move by move, its cost
  summed run by run, and the span the rule names (of the optimal ends the one on the highest diagonal, then of the
  starts an optimum ending there leaves from the one on the highest diagonal; Ties.RIGHT the same over both
  sequences reversed, each sequence's free letters swapped end for end: reference tests/test_alignment.mojo:
  1983-2020, 2054-2110) taken from that enumeration, which shares nothing with the oracle's recurrence;
- the CIGAR spells the span and costs the optimum.
"""

import itertools
import random

import pytest

from oracle import oracle as o

ALL = 1 << 40
AFFINE = o.Costs.affine(4, 6, 2)
EDIT = o.Costs.edit()


def test_reference_answers():
    # tests/test_c_api.cpp:111-121
    assert o.ends_align("TTTTACGTACGTTTTT", "ACGTACGT", AFFINE, (ALL, ALL, 0, 0)) == (0, "8=", (4, 12, 0, 8))
    cost, _, spans = o.ends_align("TTTTACGTACGTTTTT", "ACGTCGT", EDIT, (ALL, ALL, 0, 0))
    assert cost == 1 and spans[0] == 4
    assert o.ends_cost("TTTTACGTACGTTTTT", "ACGTACGT", AFFINE, (0, 0, 0, 0))[0] > 0
    assert o.ends_cost("TTTTACGTACGTTTTT", "ACGTACGT", AFFINE, (ALL, ALL, 0, 0))[0] == 0
    assert o.ends_cost("ACGTACGTTTTT", "ACGTACGT", AFFINE, (0, ALL, 0, 0))[0] == 0  # prefix
    assert o.ends_cost("TTTTACGTACGT", "ACGTACGT", AFFINE, (ALL, 0, 0, 0))[0] == 0  # suffix
    assert o.ends_cost("TTTTACGTACGT", "ACGTACGT", EDIT, (ALL, 0, 0, 0))[0] == 0
    found = o.ends_align("TTTTTACGTACGT", "ACGTACGTGGGGG", AFFINE, (5, 0, 0, 5))
    assert found[:2] == (0, "8=") and found[2][0] == 5 and found[2][3] == 8
    # tests/test_cli.sh:28-29: a read placed in a reference, infix.
    chrom = "TTTTTTTTTTACGTACGTTTGCAGGATCCAGTTTTTTTTTT"
    assert o.ends_align(chrom, "ACGTACGTTTGCAGGATCCAG", AFFINE, (ALL, ALL, 0, 0)) == (0, "21=", (10, 31, 0, 21))
    # With no free letters it is the global alignment.
    assert o.ends_align("ACGTACGTTTGCA", "ACGTCGTTTTGCA", EDIT, (0, 0, 0, 0)) == (2, "4=1D2=1I6=", (0, 13, 0, 13))


def path_cost(ref, qry, r0, c0, moves, costs):
    """A path's cost from (r0, c0): letters aligned, and each maximal gap run at its cheaper piece."""
    total, r, c, i = 0, r0, c0, 0
    while i < len(moves):
        move = moves[i]
        if move == "M":
            total += 0 if ref[r] == qry[c] else costs.mismatch
            r, c = r + 1, c + 1
            i += 1
            continue
        run = 1
        while i + run < len(moves) and moves[i + run] == move:
            run += 1
        if move == "D":
            one = costs.del_open + run * costs.del_ext
            two = costs.del_open2 + run * costs.del_ext2 if costs.pieces == 2 else one
            r += run
        else:
            one = costs.ins_open + run * costs.ins_ext
            two = costs.ins_open2 + run * costs.ins_ext2 if costs.pieces == 2 else one
            c += run
        total += min(one, two)
        i += run
    return total


def paths(dr, dc):
    """Every move sequence taking dr reference letters and dc query letters."""
    if dr == 0 and dc == 0:
        yield ""
        return
    if dr > 0 and dc > 0:
        for rest in paths(dr - 1, dc - 1):
            yield "M" + rest
    if dr > 0:
        for rest in paths(dr - 1, dc):
            yield "D" + rest
    if dc > 0:
        for rest in paths(dr, dc - 1):
            yield "I" + rest


def brute_left(ref, qry, costs, ends):
    n, m = len(ref), len(qry)
    sr, er, sc, ec = min(ends[0], n), min(ends[1], n), min(ends[2], m), min(ends[3], m)
    starts = [(r, 0) for r in range(sr + 1)] + [(0, c) for c in range(1, sc + 1)]
    finishes = [(n, c) for c in range(m - ec, m + 1)] + [(r, m) for r in range(n - er, n)]
    best = {}
    for (r0, c0), (r1, c1) in itertools.product(starts, finishes):
        if r1 < r0 or c1 < c0:
            continue
        best[(r0, c0, r1, c1)] = min(path_cost(ref, qry, r0, c0, p, costs) for p in paths(r1 - r0, c1 - c0))
    least = min(best.values())
    optimal = [k for k, v in best.items() if v == least]
    end = max(((k[2], k[3]) for k in optimal), key=lambda e: e[0] - e[1])
    start = max(((k[0], k[1]) for k in optimal if (k[2], k[3]) == end), key=lambda s: s[0] - s[1])
    return least, (start[0], end[0], start[1], end[1])


def brute(ref, qry, costs, ends, right):
    if not right:
        return brute_left(ref, qry, costs, ends)
    n, m = len(ref), len(qry)
    cost, s = brute_left(ref[::-1], qry[::-1], costs, (ends[1], ends[0], ends[3], ends[2]))
    return cost, (n - s[1], n - s[0], m - s[3], m - s[2])


COSTS = [AFFINE, EDIT, o.Costs.affine(3, 1, 2), o.Costs(4, 2, 1, 5, 3), o.Costs.two_piece(2, 2, 2, 4, 1)]


@pytest.mark.parametrize("costs", COSTS, ids=["affine", "edit", "linear-ish", "asymmetric", "two-piece"])
def test_span_rule_by_enumeration(costs):
    rng = random.Random(23)
    for trial in range(120):
        alphabet = "AC" if trial % 2 else "ACG"
        ref = "".join(rng.choice(alphabet) for _ in range(rng.randint(0, 4)))
        qry = "".join(rng.choice(alphabet) for _ in range(rng.randint(0, 4)))
        ends = tuple(rng.choice([0, 1, 2, ALL]) for _ in range(4))
        for right in (False, True):
            cost, cigar, spans = o.ends_align(ref, qry, costs, ends, right_ties=right)
            want_cost, want_spans = brute(ref, qry, costs, ends, right)
            assert (cost, spans) == (want_cost, want_spans), (ref, qry, ends, right)
            assert o.ends_cost(ref, qry, costs, ends)[0] == want_cost
            # The CIGAR spells the span and costs the optimum.
            part_r, part_q = ref[spans[0]:spans[1]], qry[spans[2]:spans[3]]
            assert o.cigar_cost(part_r, part_q, costs, cigar) == cost, (ref, qry, ends, cigar)


def test_left_rule_is_the_right_rule_mirrored():
    # Ties.LEFT of a pair = Ties.RIGHT of both sequences reversed, free letters swapped, read back (the reference's
    # own check, tests/test_alignment.mojo:2113).
    rng = random.Random(29)
    for _ in range(200):
        ref = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 30)))
        qry = "".join(rng.choice("ACGT") for _ in range(rng.randint(0, 30)))
        ends = tuple(rng.choice([0, 3, 10, ALL]) for _ in range(4))
        left = o.ends_align(ref, qry, AFFINE, ends)
        mirror = o.ends_align(ref[::-1], qry[::-1], AFFINE, (ends[1], ends[0], ends[3], ends[2]), right_ties=True)
        n, m = len(ref), len(qry)
        assert left[0] == mirror[0]
        assert left[2] == (n - mirror[2][1], n - mirror[2][0], m - mirror[2][3], m - mirror[2][2])
        runs = [r for r in __import__("re").findall(r"\d+\D", mirror[1])]
        assert left[1] == "".join(reversed(runs))

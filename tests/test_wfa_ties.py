"""A second formulation of the gap-affine tie rule, pinning the oracle's DP-flag CIGARs.

The oracle (oracle/oracle.c) writes each cell's flag by the priority of reference lanes.mojo:905-923 and walks
them back (lanes.mojo:1410-1460). The reference states the same rule a second time, in its wavefront aligner:
each cost's furthest-reaching fronts record which source won them, ties going to a substitution, then a
reference letter alone (the first sequence's), then a query letter alone, the second gap piece before the first,
an extension before an opening (reference gap_affine.mojo:954-971), and the traceback follows those flags from
the corner through every kept front (gap_affine.mojo:2028-2125). Ties.LEFT is the same over both sequences
reversed. The reference claims the two give one CIGAR byte for byte (lanes.mojo:40-46).

`wfa_align` below restates that wavefront search and traceback in plain Python over costs, diagonals and
furthest columns: nothing is shared with the oracle's cell-by-cell recurrence. Tie-rich pairs (two- and three-
letter alphabets, repeats, runs) under one-piece, two-piece and deletion-priced costs must give the oracle's
cost and CIGAR under both tie rules.
"""

import random

import pytest

from oracle import oracle as o

UNREACHED = -(1 << 40)
ALIGNED, FIRST_GAP, SECOND_GAP, FIRST_GAP2, SECOND_GAP2 = 0, 1, 2, 3, 4


def along_first(layer):
    return layer in (FIRST_GAP, FIRST_GAP2)


def opened_bit(layer):
    return 1 << (2 + layer)


def wfa_align(first: bytes, second: bytes, costs: o.Costs):
    """(cost, moves right to left) of the wavefront search over `first` (columns, the reference) and
    `second` (rows, the query): moves are ALIGNED, FIRST_GAP (a reference letter alone) or SECOND_GAP."""
    columns, rows = len(first), len(second)
    x = costs.mismatch
    # Each gap layer's opening and extension (gap_affine.mojo:249-258).
    layers = [FIRST_GAP, SECOND_GAP] + ([FIRST_GAP2, SECOND_GAP2] if costs.pieces == 2 else [])
    opening = {FIRST_GAP: costs.del_open, SECOND_GAP: costs.ins_open, FIRST_GAP2: costs.del_open2,
               SECOND_GAP2: costs.ins_open2}
    extension = {FIRST_GAP: costs.del_ext, SECOND_GAP: costs.ins_ext, FIRST_GAP2: costs.del_ext2,
                 SECOND_GAP2: costs.ins_ext2}

    def slide(column, diagonal):
        while column < columns and column - diagonal < rows and first[column] == second[column - diagonal]:
            column += 1
        return column

    # fronts[s][layer][diagonal] = furthest column; flags[s][diagonal]
    fronts, flags = [], []

    def front(s, layer, diagonal):
        if s < 0 or s >= len(fronts):
            return UNREACHED
        return fronts[s][layer].get(diagonal, UNREACHED)

    end = columns - rows
    fronts.append({layer: {} for layer in [ALIGNED] + layers})
    fronts[0][ALIGNED][0] = slide(0, 0)
    flags.append({})
    s = 0
    while fronts[s][ALIGNED].get(end, UNREACHED) < columns:
        s += 1
        lows, highs = [], []
        sources = [(s - x, ALIGNED, 0)]
        for layer in layers:
            step = 1 if along_first(layer) else -1
            sources.append((s - opening[layer] - extension[layer], ALIGNED, step))
            sources.append((s - extension[layer], layer, step))
        for cost, layer, shift in sources:
            if 0 <= cost < len(fronts) and fronts[cost][layer]:
                lows.append(min(fronts[cost][layer]) + shift)
                highs.append(max(fronts[cost][layer]) + shift)
        current = {layer: {} for layer in [ALIGNED] + layers}
        current_flags = {}
        if lows:
            low, high = max(min(lows), -rows), min(max(highs), columns)
            for k in range(low, high + 1):
                same = front(s - x, ALIGNED, k)
                substituted = same + 1 if (same >= 0 and same < columns and same - k < rows) else UNREACHED
                gap, won_opened = {}, {}
                for layer in layers:
                    o_e = opening[layer] + extension[layer]
                    e = extension[layer]
                    if along_first(layer):  # from the diagonal below, one column on
                        opened = front(s - o_e, ALIGNED, k - 1)
                        extended = front(s - e, layer, k - 1)
                        source = max(opened, extended)
                        gap[layer] = source + 1 if 0 <= source < columns else UNREACHED
                    else:  # from the diagonal above, the same column, one row on
                        opened = front(s - o_e, ALIGNED, k + 1)
                        extended = front(s - e, layer, k + 1)
                        source = max(opened, extended)
                        gap[layer] = source if source >= 0 and source - k <= rows else UNREACHED
                    won_opened[layer] = opened > extended
                gapped = max(gap.values())
                best = max(substituted, gapped)
                if best < 0:
                    continue
                current[ALIGNED][k] = slide(best, k)
                for layer in layers:
                    if gap[layer] >= 0:
                        current[layer][k] = gap[layer]
                # The flag (gap_affine.mojo:954-971).
                deletion, insertion = FIRST_GAP, SECOND_GAP
                deleted, inserted = gap[FIRST_GAP], gap[SECOND_GAP]
                bits = 0
                for layer in layers:
                    if won_opened[layer]:
                        bits |= opened_bit(layer)
                if costs.pieces == 2:
                    if gap[FIRST_GAP2] >= gap[FIRST_GAP]:
                        deletion = FIRST_GAP2
                    if gap[SECOND_GAP2] >= gap[SECOND_GAP]:
                        insertion = SECOND_GAP2
                    deleted = max(deleted, gap[FIRST_GAP2])
                    inserted = max(inserted, gap[SECOND_GAP2])
                entry = deletion if deleted >= inserted else insertion
                if substituted >= gapped:
                    entry = ALIGNED
                current_flags[k] = entry | bits
        fronts.append(current)
        flags.append(current_flags)
        if s > 4 * (columns + rows + 2) * max(x, 1) + 1000:
            raise RuntimeError("the search did not reach the corner")

    def flag(cost, diagonal):
        return flags[cost].get(diagonal, 0) if 0 <= cost < len(flags) else 0

    # The traceback (gap_affine.mojo:2028-2125), moves right to left.
    moves = []
    spent, current, k, c = s, ALIGNED, end, columns
    while True:
        if current == ALIGNED:
            row = c - k
            if c == 0:
                moves += [SECOND_GAP] * row
                break
            if row == 0:
                moves += [FIRST_GAP] * c
                break
            if spent == 0:
                start = max(k, 0)
                moves += [ALIGNED] * (c - start)
                c = start
                continue
            source = flag(spent, k) & 7
            if source == ALIGNED:
                entry = front(spent - x, ALIGNED, k) + 1
            else:
                o_ = opening[source]
                e = extension[source]
                step = 1 if along_first(source) else -1
                gap_cost, gap_diagonal, letters = spent, k, 0
                while gap_cost > 0 and flag(gap_cost, gap_diagonal) & opened_bit(source) == 0:
                    gap_cost -= e
                    gap_diagonal -= step
                    letters += 1
                if along_first(source):
                    entry = letters if gap_cost == 0 else front(gap_cost - o_ - e, ALIGNED, gap_diagonal - 1) + \
                        letters + 1
                else:
                    entry = 0 if gap_cost == 0 else front(gap_cost - o_ - e, ALIGNED, gap_diagonal + 1)
            if entry <= c:
                moves += [ALIGNED] * (c - entry)
                c = entry
            if source == ALIGNED:
                moves.append(ALIGNED)
                c -= 1
                spent -= x
            else:
                current = source
        else:
            was_opened = flag(spent, k) & opened_bit(current) != 0
            if along_first(current):
                moves.append(FIRST_GAP)
                c -= 1
                k -= 1
                if c == 0:
                    moves += [SECOND_GAP] * (-k)
                    break
            else:
                moves.append(SECOND_GAP)
                k += 1
                if c - k == 0:
                    moves += [FIRST_GAP] * c
                    break
            if was_opened:
                spent -= opening[current] + extension[current]
                current = ALIGNED
            else:
                spent -= extension[current]
    return s, moves


def wfa_cigar(ref: str, qry: str, costs: o.Costs, right_ties: bool):
    """(cost, CIGAR) by the wavefront rule: forward for Ties.RIGHT, over both sequences reversed for LEFT."""
    a, b = ref.encode(), qry.encode()
    if right_ties:
        cost, moves = wfa_align(a, b, costs)
        ordered = moves[::-1]
    else:
        cost, moves = wfa_align(a[::-1], b[::-1], costs)
        ordered = moves  # right to left over the reversed pair is left to right over the pair
    out, r, q = [], 0, 0
    for move in ordered:
        if move == FIRST_GAP:
            op = "D"
            r += 1
        elif move == SECOND_GAP:
            op = "I"
            q += 1
        else:
            op = "=" if ref[r] == qry[q] else "X"
            r += 1
            q += 1
        if out and out[-1][1] == op:
            out[-1][0] += 1
        else:
            out.append([1, op])
    assert (r, q) == (len(ref), len(qry))
    return cost, "".join(f"{n}{op}" for n, op in out)


def price(cigar: str, costs: o.Costs) -> int:
    """What a CIGAR's edits cost, each gap run at its cheaper piece."""
    import re
    total = 0
    for count, op in re.findall(r"(\d+)(\D)", cigar):
        n = int(count)
        if op == "X":
            total += n * costs.mismatch
        elif op in "DI":
            o1, e1, o2, e2 = ((costs.del_open, costs.del_ext, costs.del_open2, costs.del_ext2) if op == "D" else
                              (costs.ins_open, costs.ins_ext, costs.ins_open2, costs.ins_ext2))
            run = o1 + n * e1
            if costs.pieces == 2:
                run = min(run, o2 + n * e2)
            total += run
    return total


def tie_rich_pairs(count, seed):
    rng = random.Random(seed)
    out = []
    for _ in range(count):
        alphabet = rng.choice(["AC", "ACG", "A", "ACGT"])
        kind = rng.random()
        if kind < 0.4:
            unit = "".join(rng.choice(alphabet) for _ in range(rng.randint(1, 4)))
            ref = unit * rng.randint(1, 12)
        else:
            ref = "".join(rng.choice(alphabet) for _ in range(rng.randint(0, 36)))
        qry = list(ref)
        for _ in range(rng.randint(0, 6)):
            event = rng.random()
            at = rng.randint(0, len(qry))
            if event < 0.35 and at < len(qry):
                del qry[at:at + rng.randint(1, 8)]
            elif event < 0.7:
                qry[at:at] = [rng.choice(alphabet) for _ in range(rng.randint(1, 8))]
            elif at < len(qry):
                qry[at] = rng.choice(alphabet)
        out.append((ref, "".join(qry)))
    return out


COST_SETS = {
    "affine(4,6,2)": o.Costs.affine(4, 6, 2),
    "two_piece(4,6,2,24,1)": o.Costs.two_piece(4, 6, 2, 24, 1),
    "affine(4,6,2).with_deletions(3,1)": o.Costs.affine(4, 6, 2).with_deletions(3, 1),
    "affine(2,2,1).with_deletions(8,3)": o.Costs.affine(2, 2, 1).with_deletions(8, 3),
    "two_piece(3,4,2,9,1)": o.Costs.two_piece(3, 4, 2, 9, 1),
    # The pieces cost alike at two and three letters: ties between them are common.
    "two_piece(2,2,2,4,1)": o.Costs.two_piece(2, 2, 2, 4, 1),
    "two_piece(3,1,2,4,1)": o.Costs.two_piece(3, 1, 2, 4, 1),
    "unit": o.Costs.edit(),
}


def test_wavefront_rule_reproduces_the_reference_literals():
    # Its own check first: the reference's literals under both rules (tests/test_alignment.mojo:2162-2163,2199,
    # tests/test_c_api.cpp:69-75,95,127-128,139,155-156).
    two = o.Costs.two_piece(4, 6, 2, 24, 1)
    cases = [
        ("ACGTTTTACG", "ACGTTTACG", o.Costs.affine(4, 6, 2), False, (8, "3=1D6=")),
        ("ACGTTTTACG", "ACGTTTACG", o.Costs.affine(4, 6, 2), True, (8, "6=1D3=")),
        ("ACGTTTTACG", "ACGTTTACG", o.Costs.edit(), False, (1, "3=1D6=")),
        ("ACGTACGTTTGCA", "ACGTCGTTTTGCA", o.Costs.edit(), False, (2, "4=1D2=1I6=")),
        ("ACGTACGTTTGCA", "ACGTCGTTTTGCA", o.Costs.edit(), True, (2, "4=1D5=1I3=")),
        ("ACGTACGTTTGCA", "ACGTCGTTTTGCA", o.Costs.affine(4, 6, 2), False, (12, "4=3X6=")),
        ("GATTACAGCTTGCA" + "C" * 30 + "TGGACCATGAGTCATTGACCAGTCGATC",
         "GATTACAGCTTGCATGGACCATGAGTCAGTTGACCAGTCGATC", two, False, (62, "14=30D14=1I14=")),
        ("ACGACGT", "ACGTACGT", o.Costs.affine(4, 2, 1).with_deletions(20, 5), False, (3, "3=1I4=")),
        ("ACGTTTGCAAC", "ACGTGCATAC", o.Costs.edit(), False, (3, "3=2D4=1I2=")),
        ("", "ACG", o.Costs.edit(), False, (3, "3I")),
        ("ACGT", "", o.Costs.affine(4, 6, 2), False, (14, "4D")),
        ("", "", o.Costs.edit(), False, (0, "")),
    ]
    for ref, qry, costs, right, expected in cases:
        assert wfa_cigar(ref, qry, costs, right) == expected, (ref, qry, right)


@pytest.mark.parametrize("name", sorted(COST_SETS))
def test_dp_flag_rule_matches_the_wavefront_rule(name):
    costs = COST_SETS[name]
    pairs = tie_rich_pairs(420, seed=sorted(COST_SETS).index(name) + 7)
    for right in (False, True):
        for ref, qry in pairs:
            want = wfa_cigar(ref, qry, costs, right)
            got = o.costs_align(ref, qry, costs, right_ties=right)
            assert got == want, (name, right, ref, qry, got, want)
            assert price(got[1], costs) == got[0]

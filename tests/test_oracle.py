"""Pins the oracle (oracle/oracle.c) before anything is checked against it.

Three kinds of evidence, none of which retypes the oracle's recurrence:
- the reference's own known answers, each cited;
- brute force: every alignment of small pairs enumerated and rescored by a rescorer that shares no code with
  the sweeps (the reference's ``best_enumerated``/``brute_optimum`` pattern, tests/test_alignment.mojo:189-221);
- a second formulation from another reference file: at unit costs the bit-parallel traceback's memoryless rule
  (traceback.mojo:400-575), which must spell the same CIGAR as the lanes' flag rule (lanes.mojo:40-46).
"""

import functools
import itertools
import random

import pytest

from oracle import oracle as o

AFFINE = o.Costs.affine(4, 6, 2)
EDIT = o.Costs.edit()
TWO = o.Costs.two_piece(4, 6, 2, 24, 1)
DNA = o.uniform_scoring(2, -4)


# ------------------------------------------------------------------ known answers


@pytest.mark.parametrize(
    "ref,qry,costs,right,eqx,expected",
    [
        # tests/test_c_api.cpp:69-75, tests/test_python.py:70-72
        ("ACGTACGTTTGCA", "ACGTCGTTTTGCA", EDIT, False, True, (2, "4=1D2=1I6=")),
        ("ACGTACGTTTGCA", "ACGTCGTTTTGCA", EDIT, False, False, (2, "4M1D2M1I6M")),
        # tests/test_c_api.cpp:155-156 (WFA2-lib's rule)
        ("ACGTACGTTTGCA", "ACGTCGTTTTGCA", EDIT, True, True, (2, "4=1D5=1I3=")),
        # tests/test_alignment.mojo:2162-2163, tests/test_python.py:73-74
        ("ACGTTTTACG", "ACGTTTACG", AFFINE, False, True, (8, "3=1D6=")),
        ("ACGTTTTACG", "ACGTTTACG", AFFINE, True, True, (8, "6=1D3=")),
        # tests/test_alignment.mojo:2199 (unit costs)
        ("ACGTTTTACG", "ACGTTTACG", EDIT, False, True, (1, "3=1D6=")),
        # tests/test_c_api.cpp:95
        ("ACGTACGTTTGCA", "ACGTCGTTTTGCA", AFFINE, False, True, (12, "4=3X6=")),
        # tests/test_c_api.cpp:78-81, 96-97
        ("", "ACG", EDIT, False, True, (3, "3I")),
        ("ACG", "", EDIT, False, True, (3, "3D")),
        ("", "", EDIT, False, True, (0, "")),
        ("ACGT", "", AFFINE, False, True, (14, "4D")),
        # tests/test_c_api.cpp:84
        ("ACGTNNAC", "ACGTNNAC", EDIT, False, True, (0, "8=")),
        # tests/test_c_api.cpp:127-128 (two-piece)
        ("GATTACAGCTTGCA" + "C" * 30 + "TGGACCATGAGTCATTGACCAGTCGATC",
         "GATTACAGCTTGCATGGACCATGAGTCAGTTGACCAGTCGATC", TWO, False, True, (62, "14=30D14=1I14=")),
        # tests/test_alignment.mojo:1185 (letters outside ACGT match themselves)
        ("acgu", "acgu", AFFINE, False, True, (0, "4=")),
        # tests/test_alignment.mojo:2794-2795
        ("ACGTTTGCAAC", "ACGTGCATAC", EDIT, False, False, (3, "3M2D4M1I2M")),
        # tests/test_c_api.cpp:139 (deletions dearer than insertions)
        ("ACGACGT", "ACGTACGT", o.Costs.affine(4, 2, 1).with_deletions(20, 5), False, True, (3, "3=1I4=")),
    ],
)
def test_known_alignments(ref, qry, costs, right, eqx, expected):
    assert o.costs_align(ref, qry, costs, right_ties=right, eqx=eqx) == expected


@pytest.mark.parametrize(
    "ref,qry,costs,band,expected",
    [
        ("ACGTACGTTTGCA", "ACGTCGTTTTGCA", EDIT, None, 2),  # tests/test_python.py:70
        ("ACGTN", "ACGTA", EDIT, None, 1),  # tests/test_c_api.cpp:85
        ("NRYKMSW", "NRYKMSA", EDIT, None, 1),  # tests/test_c_api.cpp:87
        ("ACGTACGTTTGCA", "ACGTCGTTTTGCA", AFFINE, None, 12),  # tests/test_c_api.cpp:103
        ("AAAACCCC", "CCCCAAAA", AFFINE, (0, 0), 32),  # tests/test_python.py:79
        ("AAAACCCC", "CCCCAAAA", EDIT, (0, 0), 8),  # tests/test_c_api.cpp:145
        ("ACGACGT", "ACGTACGT", AFFINE.with_deletions(20, 5), None, 8),  # tests/test_python.py:85
        ("ACGTACGT", "ACGACGT", AFFINE.with_deletions(20, 5), None, 25),  # tests/test_python.py:86
        ("ACGTACGT", "ACGACGT", o.Costs.affine(4, 2, 1).with_deletions(20, 5), None, 25),  # test_c_api.cpp:137
        ("GATTACAGCTTGCA" + "C" * 30 + "TGGACCATGAGTCATTGACCAGTCGATC",
         "GATTACAGCTTGCATGGACCATGAGTCAGTTGACCAGTCGATC", AFFINE, None, 74),  # tests/test_c_api.cpp:129
        ("GATTACAGCTTGCA" + "C" * 30 + "TGGACCATGAGTCA", "GATTACAGCTTGCATGGACCATGAGTCA", TWO, None,
         54),  # tests/test_python.py:84
        ("ACGT", "AC", AFFINE, (-1, 1), o.OUTSIDE),  # tests/test_c_api.cpp:147-150
    ],
)
def test_known_costs(ref, qry, costs, band, expected):
    band = band or (o.NO_BAND_LOW, o.NO_BAND_HIGH)
    assert o.costs_cost(ref, qry, costs, band) == expected


def test_known_band_alignment():
    # tests/test_c_api.cpp:143: a band of the one diagonal, substitutions only.
    assert o.costs_align("ACGTACGT", "ACGAACGT", AFFINE, (0, 0)) == (4, "3=1X4=")


@pytest.mark.parametrize(
    "ref,qry,expected",
    [  # tests/test_alignment.mojo:343-365 (known Levenshtein distances)
        ("ACGT", "ACGT", 0), ("ACGT", "AGT", 1), ("ACGT", "", 4), ("", "ACG", 3), ("AAAA", "TTTT", 4),
        ("ACGTACGT", "TACGTACG", 2), ("ACGTACGTTTGCA", "ACGTCGTTTTGCA", 2),
    ],
)
def test_known_distances(ref, qry, expected):
    assert o.edit_distance(ref, qry) == expected
    assert o.myers_distance(ref, qry) == expected
    assert o.costs_cost(ref, qry, EDIT) == expected


def test_known_scoring():
    # tests/test_alignment.mojo:294-340 (hand-computed global and local, Scoring.dna()).
    assert o.scoring_global_align("ACGTACGT", "ACGTACGT", DNA) == (16, "8=")
    assert o.scoring_global_align("ACGTACGT", "ACGTTCGT", DNA) == (10, "4=1X3=")
    assert o.scoring_global_align("ACGTACGT", "ACGTTCGT", DNA, eqx=False) == (10, "8M")
    assert o.scoring_global_align("ACGTTGCAGGGCATGACGT", "ACGTTGCACATGACGT", DNA) == (22, "8=3D8=")
    assert o.scoring_global_score("AAAA", "", DNA) == -6 - 3 * 2
    assert o.scoring_global_score("", "", DNA) == 0
    score, cigar, spans = o.scoring_local_align("GGGGACGTACGTGGGG", "CCCCACGTACGTCCCC", DNA)
    assert (score, cigar, spans[0], spans[3]) == (16, "8=", 4, 12)
    # tests/test_python.py:98-99, 103-104
    assert o.scoring_local_align("TTTTACGTACGTTTTT", "ACGTACGT", DNA)[0] == 16
    blosum_like = o.Scoring("ACGT", (5, -1, -2, -1, -1, 5, -3, -2, -2, -3, 5, -1, -1, -2, -1, 5), -6, -1)
    assert o.scoring_global_align("ACGTAC", "ACGTAC", blosum_like)[0] == 30


# ------------------------------------------------------------------ brute force


def alignments(a: str, b: str):
    """Every global alignment of a and b as two gapped rows."""

    @functools.lru_cache(maxsize=None)
    def go(i, j):
        if i == len(a) and j == len(b):
            return [("", "")]
        out = []
        if i < len(a) and j < len(b):
            out += [(a[i] + t, b[j] + u) for t, u in go(i + 1, j + 1)]
        if i < len(a):
            out += [(a[i] + t, "-" + u) for t, u in go(i + 1, j)]
        if j < len(b):
            out += [("-" + t, b[j] + u) for t, u in go(i, j + 1)]
        return out

    return go(0, 0)


def rows_of(cigar: str, a: str, b: str):
    """The gapped rows a CIGAR spells over a and b."""
    top, bottom, i, j, num = [], [], 0, 0, ""
    for ch in cigar:
        if ch.isdigit():
            num += ch
            continue
        for _ in range(int(num)):
            if ch == "D":
                top.append(a[i]); bottom.append("-"); i += 1
            elif ch == "I":
                top.append("-"); bottom.append(b[j]); j += 1
            else:
                assert ch in "=XM"
                if ch == "=":
                    assert a[i] == b[j]
                if ch == "X":
                    assert a[i] != b[j]
                top.append(a[i]); bottom.append(b[j]); i += 1; j += 1
        num = ""
    assert i == len(a) and j == len(b)
    return "".join(top), "".join(bottom)


def path_diagonals(top: str, bottom: str):
    """The library diagonals (reference letters less query letters) a pair of rows visits, origin included."""
    r = c = 0
    seen = {0}
    for x, y in zip(top, bottom):
        r += x != "-"
        c += y != "-"
        seen.add(r - c)
    return seen


def words(alphabet, max_len):
    for n in range(max_len + 1):
        for t in itertools.product(alphabet, repeat=n):
            yield "".join(t)


COSTS = [EDIT, AFFINE, TWO, o.Costs.affine(4, 2, 1).with_deletions(9, 1), o.Costs.affine(1, 0, 1)]


@pytest.mark.parametrize("costs", COSTS)
def test_costs_match_enumeration(costs):
    """The least cost equals the cheapest enumerated alignment; the traced CIGAR is optimal and rebuilds both."""
    pairs = [(a, b) for a in words("AC", 3) for b in words("AC", 3)]
    for a, b in pairs:
        best = min(o.rows_cost(t, u, costs) for t, u in alignments(a, b))
        assert o.costs_cost(a, b, costs) == best
        for right in (False, True):
            cost, cigar = o.costs_align(a, b, costs, right_ties=right)
            assert cost == best
            top, bottom = rows_of(cigar, a, b)
            assert o.rows_cost(top, bottom, costs) == best


def test_band_matches_enumeration():
    """A band's cost is the cheapest alignment whose every cell lies inside it (modes.mojo:429-437)."""
    rng = random.Random(3)
    for _ in range(300):
        a = "".join(rng.choice("ACG") for _ in range(rng.randint(0, 4)))
        b = "".join(rng.choice("ACG") for _ in range(rng.randint(0, 4)))
        low = rng.randint(-3, 1)
        high = rng.randint(low, 3)
        inside = [(t, u) for t, u in alignments(a, b) if path_diagonals(t, u) <= set(range(low, high + 1))]
        want = min((o.rows_cost(t, u, AFFINE) for t, u in inside), default=o.OUTSIDE)
        assert o.costs_cost(a, b, AFFINE, (low, high)) == want
        cost, cigar = o.costs_align(a, b, AFFINE, (low, high))
        assert cost == want
        if cigar is not None:
            top, bottom = rows_of(cigar, a, b)
            assert path_diagonals(top, bottom) <= set(range(low, high + 1))


def scoring_rescore(top, bottom, scoring):
    """What two gapped rows earn under a Scoring: the table for aligned letters, opening + k*extension a gap run."""
    index = {ch: i for i, ch in enumerate(scoring.alphabet)}
    letters = len(scoring.alphabet)
    total, run, length = 0, None, 0
    for x, y in zip(top, bottom):
        kind = "D" if y == "-" else ("I" if x == "-" else None)
        if kind != run and run is not None:
            total += scoring.opening + length * scoring.extension
            length = 0
        run = kind
        if kind is None:
            total += scoring.table[index[x] * letters + index[y]]
        else:
            length += 1
    if run is not None:
        total += scoring.opening + length * scoring.extension
    return total


SCORINGS = [
    # tests/test_alignment.mojo:109-119 (the five scoring regimes)
    o.uniform_scoring(5, -4, -19, -1), o.uniform_scoring(2, -1, -1, -1), o.uniform_scoring(0, -1, 0, -1),
    o.uniform_scoring(1, -1, -5, 0), DNA,
    o.Scoring("ACGT", (5, -1, -2, -1, -1, 5, -3, -2, -2, -3, 5, -1, -1, -2, -1, 5), -6, -1),
]


@pytest.mark.parametrize("scoring", SCORINGS)
def test_scoring_global_matches_enumeration(scoring):
    for a in words("ACG", 3):
        for b in words("AC", 3):
            best = max(scoring_rescore(t, u, scoring) for t, u in alignments(a, b))
            assert o.scoring_global_score(a, b, scoring) == best
            score, cigar = o.scoring_global_align(a, b, scoring)
            assert score == best
            top, bottom = rows_of(cigar, a, b)
            assert scoring_rescore(top, bottom, scoring) == best


def global_best(a, b, scoring):
    return max(scoring_rescore(t, u, scoring) for t, u in alignments(a, b))


@pytest.mark.parametrize("scoring", SCORINGS)
def test_scoring_local_matches_brute_force(scoring):
    """Local score, end, start and spans by brute force over every pair of substrings (the end: best score,
    then the latest anti-diagonal, then the furthest reference letter, scored.mojo:81-84; the start: the first
    cell by anti-diagonal back from the end, then row, reaching the score, vector_score.mojo:37-60)."""
    rng = random.Random(5)
    for _ in range(60):
        a = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 6)))
        b = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 5)))
        n, m = len(a), len(b)
        ends = {}
        for er in range(1, n + 1):
            for ec in range(1, m + 1):
                ends[(er, ec)] = max(global_best(a[sr:er], b[sc:ec], scoring)
                                     for sr in range(er) for sc in range(ec))
        best = max([0] + list(ends.values()))
        score, er, ec = o.scoring_local_end(a, b, scoring)
        assert score == best
        if best <= 0:
            continue
        want_end = max((v, r + c, r) for (r, c), v in ends.items() if v == best)
        assert (er, ec) == (want_end[2], want_end[1] - want_end[2])
        starts = sorted((i + j, i, er - i, ec - j) for i in range(1, er + 1) for j in range(1, ec + 1)
                        if global_best(a[er - i:er], b[ec - j:ec], scoring) >= best)
        sr, sc = starts[0][2], starts[0][3]
        got, cigar, spans = o.scoring_local_align(a, b, scoring)
        assert got == best
        top, bottom = rows_of(cigar, a[sr:er], b[sc:ec])
        assert scoring_rescore(top, bottom, scoring) == best
        assert spans == (a.rfind(a[sr:er]), a.rfind(a[sr:er]) + er - sr, b.rfind(b[sc:ec]), b.rfind(b[sc:ec]) + ec - sc)


# ------------------------------------------------------------------ a second formulation of the tie rule


def memoryless_unit_cigar(a: str, b: str, left: bool = True) -> str:
    """The bit-parallel traceback's rule at unit costs (traceback.mojo:400-575): from the corner, a substitution
    when one is optimal, else a reference letter alone, else a query letter alone, and a match only when no
    edit is; Ties.LEFT over both sequences reversed (edit_distance.mojo:295-297)."""
    x, y = (a[::-1], b[::-1]) if left else (a, b)
    n, m = len(x), len(y)
    h = [[0] * (m + 1) for _ in range(n + 1)]
    for c in range(m + 1):
        h[0][c] = c
    for r in range(1, n + 1):
        h[r][0] = r
        for c in range(1, m + 1):
            h[r][c] = min(h[r - 1][c - 1] + (x[r - 1] != y[c - 1]), h[r - 1][c] + 1, h[r][c - 1] + 1)
    r, c, moves = n, m, []
    while r > 0 or c > 0:
        cur = h[r][c]
        if r > 0 and c > 0 and x[r - 1] != y[c - 1] and h[r - 1][c - 1] == cur - 1:
            moves.append("M"); r -= 1; c -= 1
        elif r > 0 and h[r - 1][c] == cur - 1:
            moves.append("D"); r -= 1
        elif c > 0 and h[r][c - 1] == cur - 1:
            moves.append("I"); c -= 1
        else:
            moves.append("M"); r -= 1; c -= 1
    if not left:
        moves.reverse()
    out, i, j, runs = [], 0, 0, []
    for mv in moves:
        op = mv
        if mv == "M":
            op = "=" if a[i] == b[j] else "X"
            i += 1; j += 1
        elif mv == "D":
            i += 1
        else:
            j += 1
        if runs and runs[-1][0] == op:
            runs[-1][1] += 1
        else:
            runs.append([op, 1])
    return "".join(f"{k}{op}" for op, k in runs)


def test_unit_tie_rule_matches_bit_parallel_rule():
    rng = random.Random(11)
    for trial in range(400):
        alpha = ["AC", "ACG", "ACGT"][trial % 3]
        a = "".join(rng.choice(alpha) for _ in range(rng.randint(0, 40)))
        b = list(a)
        for _ in range(rng.randint(0, 8)):
            k = rng.randrange(3)
            pos = rng.randint(0, len(b))
            if k == 0 and pos < len(b):
                b[pos] = rng.choice(alpha)
            elif k == 1:
                b.insert(pos, rng.choice(alpha))
            elif pos < len(b):
                del b[pos]
        b = "".join(b)
        for left in (True, False):
            assert o.costs_align(a, b, EDIT, right_ties=not left)[1] == memoryless_unit_cigar(a, b, left)


def test_myers_matches_levenshtein():
    rng = random.Random(13)
    for trial in range(200):
        n = rng.choice([0, 1, 5, 63, 64, 65, 130, 200])
        a = "".join(rng.choice("ACGTN") for _ in range(n))
        b = "".join(rng.choice("ACGTN") for _ in range(rng.randint(0, 220)))
        assert o.myers_distance(a, b) == o.edit_distance(a, b)


# ------------------------------------------------------------------ the band certificate


def test_off_band_is_a_lower_bound_and_tight():
    """off_band (lanes.mojo:333-345) bounds every path visiting the diagonal, and a path that runs straight out
    and back reaches it."""
    assert o.off_band(AFFINE, 3, 0) == 2 * (6 + 3 * 2)
    assert o.off_band(AFFINE, -2, 1) == (6 + 2 * 2) + (6 + 3 * 2)
    assert o.off_band(AFFINE, 1, 2) == 0
    rng = random.Random(17)
    for _ in range(200):
        a = "".join(rng.choice("AC") for _ in range(rng.randint(1, 4)))
        b = "".join(rng.choice("AC") for _ in range(rng.randint(1, 4)))
        end = len(b) - len(a)
        for k in range(-len(a), len(b) + 1):
            visiting = [o.rows_cost(t, u, AFFINE) for t, u in alignments(a, b) if -k in path_diagonals(t, u)]
            if visiting:
                assert min(visiting) >= o.off_band(AFFINE, k, end)


@pytest.mark.parametrize("costs", [AFFINE, EDIT, o.Costs.affine(4, 2, 1).with_deletions(9, 1)])
def test_proven_band_traces_the_whole_matrix_alignment(costs):
    """A band the strict certificate proves (lanes.mojo:1337-1392) holds every optimal path: its cost and its
    CIGAR equal the whole matrix's. An unproven one's widened band proves itself."""
    rng = random.Random(19)
    for trial in range(150):
        n = rng.randint(1, 60)
        a = "".join(rng.choice("ACGT") for _ in range(n))
        b = list(a)
        for _ in range(rng.randint(0, 6)):
            pos = rng.randint(0, len(b))
            k = rng.randrange(3)
            if k == 0 and pos < len(b):
                b[pos] = rng.choice("ACGT")
            elif k == 1:
                b.insert(pos, rng.choice("ACGT"))
            elif pos < len(b):
                del b[pos]
        b = "".join(b)
        end = len(b) - len(a)
        whole = o.costs_align(a, b, costs)
        low = min(0, end) - rng.randint(1, 4)
        high = max(0, end) + rng.randint(1, 4)
        for strict in (True, False):
            found = o.costs_cost(a, b, costs, (-high, -low))
            proven, wide_low, wide_high = o.proof(costs, len(a), len(b), found, low, high, strict)
            if not proven:
                low, high = wide_low, wide_high
                found = o.costs_cost(a, b, costs, (-high, -low))
                assert o.proof(costs, len(a), len(b), found, low, high, strict)[0] or (
                    low <= -len(a) and high >= len(b))
            assert found == whole[0]
            if strict:
                assert o.costs_align(a, b, costs, (-high, -low)) == whole


# ------------------------------------------------------------------ the banded-memory oracle for long pairs


def _random_pairs(count, seed, longest=60):
    rng = random.Random(seed)
    out = []
    for _ in range(count):
        alphabet = rng.choice(["AC", "ACGT", "A", "ACG"])
        ref = "".join(rng.choice(alphabet) for _ in range(rng.randint(0, longest)))
        qry = list(ref)
        for _ in range(rng.randint(0, 8)):
            at = rng.randint(0, len(qry))
            event = rng.random()
            if event < 0.3 and at < len(qry):
                del qry[at:at + rng.randint(1, 10)]
            elif event < 0.6:
                qry[at:at] = [rng.choice(alphabet) for _ in range(rng.randint(1, 10))]
            elif at < len(qry):
                qry[at] = rng.choice(alphabet)
        out.append((ref, "".join(qry)))
    return out


@pytest.mark.parametrize("costs", [AFFINE, EDIT, TWO, o.Costs.affine(4, 6, 2).with_deletions(3, 1),
                                   o.Costs.two_piece(2, 2, 2, 4, 1)])
def test_banded_memory_oracle_equals_the_whole_matrix_oracle(costs):
    # The long-pair oracle stores the strict proven band alone; on every pair it must give the whole matrix's cost
    # and CIGAR under both tie rules, and its distance the whole matrix's.
    for ref, qry in _random_pairs(500, seed=91):
        for right in (False, True):
            assert o.costs_align_proven(ref, qry, costs, right_ties=right) == \
                o.costs_align(ref, qry, costs, right_ties=right), (ref, qry, right)
        assert o.costs_distance_proven(ref, qry, costs) == o.costs_cost(ref, qry, costs)


def test_banded_memory_oracle_stores_a_band_not_the_matrix():
    rng = random.Random(92)
    ref = "".join(rng.choice("ACGT") for _ in range(3000))
    qry = ref[:1000] + ref[1010:2000] + "ACGTA" + ref[2000:]
    cost, cigar, cells = o.costs_align_proven(ref, qry, EDIT, with_cells=True)
    assert (cost, cigar) == o.costs_align(ref, qry, EDIT)
    assert cells < 3001 * 40  # the strict band is a few dozen diagonals, not 6,000


def test_cigar_checker_rejects_what_does_not_spell_the_pair():
    ref, qry = "ACGTACGTTTGCA", "ACGTCGTTTTGCA"
    assert o.cigar_cost(ref, qry, AFFINE, "4=3X6=") == 12
    assert o.cigar_cost(ref, qry, EDIT, "4=1D2=1I6=") == 2
    assert o.cigar_cost(ref, qry, TWO, "4=1D2=1I6=") == 16
    assert o.cigar_cost(ref, qry, AFFINE, "4=3X5=") == -2   # a letter short
    assert o.cigar_cost(ref, qry, AFFINE, "4=3=6=") == -3   # `=` over unequal letters
    assert o.cigar_cost(ref, qry, AFFINE, "5=2X6=") == -3   # ... and over a wrong split
    assert o.cigar_cost(ref, qry, AFFINE, "4=3X6") == -2    # a count with no operation
    assert o.cigar_cost(ref, qry, AFFINE, "4=3Q6=") == -1
    assert o.check_alignments([ref, ref], [qry, qry], AFFINE, ["4=3X6=", "4=3X6="], [12, 13]) == (1, 1)

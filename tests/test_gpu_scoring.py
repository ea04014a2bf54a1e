"""GPU parity under a substitution table (Scoring): scores and alignments from libdinara_hip.so (through the C ABI)
against the oracle, pair by pair: the score, the CIGAR Gotoh's walk spells, and a local alignment's spans.

Cases follow the reference's known answers (tests/test_alignment.mojo:294-340), its five scoring regimes
(:109-119), the local configuration's shape (a 20-letter table, a query inside a longer window), and the edges:
empty sides, one-letter and 32-letter alphabets, queries wider than one strip of 512 columns, and refusals.
"""

import json
import os
import random

import pytest

import dinara_align as da
from dinara_align import synthetic
from oracle import oracle as o

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

DNA = da.Scoring.dna()
REGIMES = [
    da.Scoring.uniform(5, -4, -19, -1), da.Scoring.uniform(2, -1, -1, -1), da.Scoring.uniform(0, -1, 0, -1),
    da.Scoring.uniform(1, -1, -5, 0), DNA,
    da.Scoring.tabulated("ACGT", (5, -1, -2, -1, -1, 5, -3, -2, -2, -3, 5, -1, -1, -2, -1, 5), -6, -1),
]


def to_oracle(s: da.Scoring) -> o.Scoring:
    return o.Scoring(s.alphabet, tuple(s.table), s.opening, s.extension)


def mutated(text, rate, rng, alphabet):
    out = []
    for ch in text:
        u = rng.random()
        if u < rate / 3:
            out.append(rng.choice(alphabet))
        elif u < 2 * rate / 3:
            out.append(rng.choice(alphabet))
            out.append(ch)
        elif u < rate:
            continue
        else:
            out.append(ch)
    return "".join(out)


def random_pairs(count, rng, lengths=(1, 70), alphabet="ACGT", rates=(0.0, 0.1, 0.3, 0.6)):
    refs, qrys = [], []
    for i in range(count):
        r = "".join(rng.choice(alphabet) for _ in range(rng.randint(*lengths)))
        refs.append(r)
        qrys.append(mutated(r, rates[i % len(rates)], rng, alphabet))
    return refs, qrys


def check_global(refs, qrys, scoring, eqx=True):
    os_ = to_oracle(scoring)
    got_scores = da.scores(refs, qrys, scoring)
    got = da.alignments(refs, qrys, scoring, eqx=eqx)
    for i, (r, q) in enumerate(zip(refs, qrys)):
        want_score, want_cigar = o.scoring_global_align(r, q, os_, eqx=eqx)
        assert got_scores[i] == o.scoring_global_score(r, q, os_), (i, r, q)
        a = got[i]
        assert (a.score, a.cigar) == (want_score, want_cigar), (i, r, q)
        assert a.cost == -want_score
        assert (a.reference_start, a.reference_end, a.query_start, a.query_end) == (0, len(r), 0, len(q))


def check_local(refs, qrys, scoring, eqx=True):
    os_ = to_oracle(scoring)
    got_scores = da.scores(refs, qrys, scoring, da.Mode.local())
    got = da.alignments(refs, qrys, scoring, da.Mode.local(), eqx=eqx)
    for i, (r, q) in enumerate(zip(refs, qrys)):
        want_score, want_cigar, spans = o.scoring_local_align(r, q, os_, eqx=eqx)
        assert got_scores[i] == want_score, (i, r, q)
        a = got[i]
        assert (a.score, a.cigar) == (want_score, want_cigar), (i, r, q)
        assert (a.reference_start, a.reference_end, a.query_start, a.query_end) == tuple(spans), (i, r, q)


def test_known_answers():
    # tests/test_alignment.mojo:320-340 under Scoring.dna().
    a = da.align("ACGTTGCAGGGCATGACGT", "ACGTTGCACATGACGT", DNA)
    assert (a.score, a.cigar) == (22, "8=3D8=")
    assert da.align("ACGTACGT", "ACGTTCGT", DNA).cigar == "4=1X3="
    assert da.align("ACGTACGT", "ACGTTCGT", DNA, eqx=False).cigar == "8M"
    assert da.score("ACGTACGT", "ACGTACGT", DNA) == 16
    local = da.align("GGGGACGTACGTGGGG", "CCCCACGTACGTCCCC", DNA, da.Mode.local())
    assert (local.score, local.cigar) == (16, "8=")
    assert (local.reference_start, local.reference_end, local.query_start, local.query_end) == (4, 12, 4, 12)


@pytest.mark.parametrize("scoring", REGIMES)
def test_random_global(scoring):
    rng = random.Random(11)
    refs, qrys = random_pairs(160, rng)
    check_global(refs, qrys, scoring)


@pytest.mark.parametrize("scoring", REGIMES)
def test_random_local(scoring):
    rng = random.Random(12)
    refs, qrys = random_pairs(160, rng)
    # Unrelated pairs too: local ends anywhere, or nothing above zero.
    extra_r, extra_q = random_pairs(40, rng, rates=(1.0,))
    check_local(refs + extra_r, qrys + extra_q, scoring)


def test_eqx_off_and_empty_sides():
    refs = ["", "ACGT", "", "A", "ACGTACGT", "TTTT"]
    qrys = ["", "", "ACG", "A", "ACGAACGT", "AAAA"]
    check_global(refs, qrys, DNA, eqx=False)
    check_local(refs, qrys, DNA, eqx=False)


def test_wide_queries_take_several_strips():
    rng = random.Random(13)
    refs, qrys = [], []
    for n, m in [(40, 600), (300, 1100), (700, 530), (5, 1500), (1200, 900)]:
        q = "".join(rng.choice("ACGT") for _ in range(m))
        refs.append(mutated(q, 0.1, rng, "ACGT")[:n] if n < m else mutated(q, 0.1, rng, "ACGT") + "AC" * (n - m))
        qrys.append(q)
    check_global(refs, qrys, DNA)
    check_local(refs, qrys, DNA)


def test_alphabets_one_and_thirty_two_letters():
    rng = random.Random(14)
    one = da.Scoring.tabulated("A", (3,), -2, -1)
    check_global(["A" * 5, "AAA", "A"], ["AAA", "A" * 9, "A"], one)
    check_local(["A" * 5, "AAA"], ["AAA", "A" * 9], one)
    letters = "ABCDEFGHIJKLMNOPQRSTUVWXYZ012345"
    table = [rng.randint(-6, 9) if a != b else rng.randint(4, 11) for a in range(32) for b in range(32)]
    wide = da.Scoring.tabulated(letters, table, -7, -2)
    refs, qrys = random_pairs(80, rng, lengths=(1, 90), alphabet=letters)
    check_global(refs, qrys, wide)
    check_local(refs, qrys, wide)


def test_refusals():
    with pytest.raises(ValueError):
        da.scores(["ACGT"], ["ACGX"], DNA)  # a letter outside the alphabet
    with pytest.raises(ValueError):
        da.scores(["ACGT"], ["ACGT"], da.Scoring.uniform(2, -4, 1, -2))  # a gap opening that rewards
    with pytest.raises(ValueError):
        da.scores(["ACGT"], ["ACGT"], da.Scoring.uniform(2, -4, -4, 1))  # an extension that rewards
    with pytest.raises(ValueError):
        da.scores(["ACGT"], ["ACGT"], da.Scoring.tabulated("AC-T", tuple([1] * 16)))  # '-' is a gap
    with pytest.raises(ValueError):
        da.alignments(["ACGT"], ["ACGT"], DNA, da.Mode.local(2))  # a match score in the mode as well
    # A pair that passes the Scoring's 32-bit reach.
    big = da.Scoring.uniform(127, -128, -2_000_000, -1_000_000)
    with pytest.raises(ValueError):
        da.scores(["ACGT" * 100], ["ACGT" * 100], big)


def test_golden_local_protein():
    with open(os.path.join(GOLDEN, "c5_local.json")) as f:
        d = json.load(f)
    scoring = da.Scoring.tabulated(d["alphabet"], d["table"], d["opening"], d["extension"])
    refs = [p[0] for p in d["pairs"]]
    qrys = [p[1] for p in d["pairs"]]
    got = da.alignments(refs, qrys, scoring, da.Mode.local())
    got_scores = da.scores(refs, qrys, scoring, da.Mode.local())
    for i, (r, q, score, cigar, spans) in enumerate(d["pairs"]):
        assert got_scores[i] == score
        a = got[i]
        assert (a.score, a.cigar) == (score, cigar), i
        assert [a.reference_start, a.reference_end, a.query_start, a.query_end] == spans, i


def test_local_protein_workload_against_oracle():
    alphabet, table = synthetic.protein_table()
    scoring = da.Scoring.tabulated(alphabet, table, -11, -1)
    refs, qrys = synthetic.local_protein(24, seed=7, query=300, window=3000)
    check_local([r.decode() for r in refs], [q.decode() for q in qrys], scoring)


def test_cigar_runs_across_chunk_edges():
    # Runs longer than 2048 moves and a run ending at move 2048 exactly: the CIGAR text is made 2048 moves at a time.
    rng = random.Random(15)
    base = "".join(rng.choice("ACGT") for _ in range(4300))
    other = {"A": "C", "C": "G", "G": "T", "T": "A"}
    refs = [base[:4300], base[:4100]]
    qrys = [base[:4300], base[:2047] + other[base[2047]] + base[2048:4100]]
    check_global(refs, qrys, DNA)
    check_global(refs, qrys, DNA, eqx=False)


def test_reach_back_ignores_columns_past_the_query():
    # A lane sweeps whole strips of columns, those past the query's end on stand-in letters. Under a large match
    # score those columns reach a local end's score before the real start does; the reach back must not stop on
    # them (its cutoff counts real columns only). Queries of 3 to 70 letters inside references of 1,000 to 1,500,
    # with deletions in the embedded copy.
    rng = random.Random(99)
    refs, qrys = [], []
    for _ in range(24):
        m = rng.randint(3, 70)
        q = "".join(rng.choice("ACGT") for _ in range(m))
        copy = "".join(ch + ("".join(rng.choice("ACGT") for _ in range(rng.randint(1, 12))) if rng.random() < 0.2
                             else "") for ch in q)
        n = rng.randint(1000, 1500)
        at = rng.randint(0, n - len(copy))
        r = "".join(rng.choice("ACGT") for _ in range(at)) + copy
        refs.append(r + "".join(rng.choice("ACGT") for _ in range(n - len(r))))
        qrys.append(q)
    for match in (100, 120, 5):
        check_local(refs, qrys, da.Scoring.uniform(match, -4, -19, -1))


def test_tall_references_go_through_the_ring():
    # A reference longer than LDS stages whole (64 KiB less the profile) is read through a ring of letters
    # (reference tests/test_alignment.mojo:3801-3810: 70 000 letters against a 16-letter query, global and local).
    scoring = da.Scoring.uniform(5, -4, -19, -1)
    tall, short = "A" * 70_000, "ACGT" * 4
    # Global: the query's four A's match, its twelve other letters mismatch, the other 69 984 letters one deletion
    # (a gap of k letters scores opening + k extension, reference scoring.mojo:202-214).
    assert da.score(tall, short, scoring) == 4 * 5 - 12 * 4 + (-19 - 69_984)
    assert da.score(tall, short, scoring, da.Mode.local()) == 5
    assert da.score(tall, "", scoring) == -19 - 70_000
    assert da.score(tall, "", scoring, da.Mode.local()) == 0
    rng = random.Random(16)
    refs, qrys = [tall], [short]
    for n, m in [(66_000, 40), (70_001, 3), (80_000, 600), (65_600, 200)]:
        q = "".join(rng.choice("ACGT") for _ in range(m))
        at = rng.randint(0, n - m)
        r = "".join(rng.choice("ACGT") for _ in range(at)) + mutated(q, 0.1, rng, "ACGT")
        refs.append(r + "".join(rng.choice("ACGT") for _ in range(n - len(r))))
        qrys.append(q)
    check_global(refs, qrys, scoring, eqx=False)
    check_local(refs, qrys, scoring)
    check_local(refs, qrys, DNA)
    # Local ends are kept as one key a column, score above a row of 17 bits here, while the scores fit the 15 bits
    # left; a match of 120 against a 600-letter query does not, and the ends are kept by compare and select.
    check_local(refs, qrys, da.Scoring.uniform(120, -4, -19, -1))

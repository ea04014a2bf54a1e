"""Writes the golden fixtures: seeded pairs shaped like each benchmark configuration with the oracle's answers.

Run from the repository root: ``python tests/golden/make_golden.py``. The oracle (oracle/oracle.c) is pinned by
tests/test_oracle.py against the reference's own known answers and brute force; these fixtures freeze its
answers so the GPU tests compare against data, and tests/test_golden.py keeps the oracle to them.

Fixtures (JSON, 256 pairs each unless noted):
- c1_edit.json: 150 bp at 2%, unit costs, distances.
- c2_affine.json: 250 bp at 5%, costs (4, 6, 2), cost and Ties.LEFT CIGAR (and Ties.RIGHT for 64).
- c3_uneven_small.json: the uneven generator clipped to 100..1500 bp (the full 20 kbp clip is too large for the
  full-matrix oracle), unit costs, distance and CIGAR.
- c4_affine_uneven_small.json: the same pairs under (4, 6, 2), distances.
- c5_local.json: 64 local alignments, 300 aa query in a 3000 aa window, the 20-letter table, gap -11/-1.
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dinara-align_amd"))

from oracle import oracle as o  # noqa: E402
from dinara_align import synthetic  # noqa: E402


def dump(name, payload):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(payload, f, separators=(",", ":"))
        f.write("\n")


def main():
    refs, qrys = synthetic.edit_short(256, seed=101)
    dump("c1_edit.json", {"costs": [1, 0, 1], "pairs": [
        [r.decode(), q.decode(), o.edit_distance(r, q)] for r, q in zip(refs, qrys)]})

    refs, qrys = synthetic.affine_reads(256, seed=102)
    affine = o.Costs.affine(4, 6, 2)
    pairs = []
    for i, (r, q) in enumerate(zip(refs, qrys)):
        cost, cigar = o.costs_align(r, q, affine)
        entry = [r.decode(), q.decode(), cost, cigar]
        if i < 64:
            entry.append(o.costs_align(r, q, affine, right_ties=True)[1])
        pairs.append(entry)
    dump("c2_affine.json", {"costs": [4, 6, 2], "pairs": pairs})

    refs, qrys = synthetic.uneven(256, seed=103, clip=(100, 1500))
    edit = o.Costs.edit()
    dump("c3_uneven_small.json", {"costs": [1, 0, 1], "pairs": [
        [r.decode(), q.decode(), *o.costs_align(r, q, edit)] for r, q in zip(refs, qrys)]})
    dump("c4_affine_uneven_small.json", {"costs": [4, 6, 2], "pairs": [
        [r.decode(), q.decode(), o.costs_cost(r, q, affine)] for r, q in zip(refs, qrys)]})

    alphabet, table = synthetic.protein_table()
    scoring = o.Scoring(alphabet, table, -11, -1)
    refs, qrys = synthetic.local_protein(64, seed=105)
    pairs = []
    for r, q in zip(refs, qrys):
        score, cigar, spans = o.scoring_local_align(r, q, scoring)
        pairs.append([r.decode(), q.decode(), score, cigar, list(spans)])
    dump("c5_local.json", {"alphabet": alphabet, "table": list(table), "opening": -11, "extension": -1,
                           "pairs": pairs})


if __name__ == "__main__":
    main()

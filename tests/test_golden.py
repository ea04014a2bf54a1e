"""The oracle against the committed golden fixtures (tests/golden/make_golden.py wrote them): the fixtures
freeze answers the pinned oracle gave, and this keeps the oracle from drifting off them."""

import json
import os

import pytest

from oracle import oracle as o

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_c1_edit_fixture():
    for ref, qry, cost in load("c1_edit.json")["pairs"]:
        assert o.edit_distance(ref, qry) == cost
        assert o.myers_distance(ref, qry) == cost


def test_c2_affine_fixture():
    affine = o.Costs.affine(4, 6, 2)
    for entry in load("c2_affine.json")["pairs"]:
        ref, qry, cost, cigar = entry[:4]
        assert o.costs_align(ref, qry, affine) == (cost, cigar)
        if len(entry) > 4:
            assert o.costs_align(ref, qry, affine, right_ties=True) == (cost, entry[4])


def test_c3_c4_fixtures():
    edit, affine = o.Costs.edit(), o.Costs.affine(4, 6, 2)
    for ref, qry, cost, cigar in load("c3_uneven_small.json")["pairs"][:48]:
        assert o.costs_align(ref, qry, edit) == (cost, cigar)
        assert o.myers_distance(ref, qry) == cost
    for ref, qry, cost in load("c4_affine_uneven_small.json")["pairs"][:48]:
        assert o.costs_cost(ref, qry, affine) == cost


def test_c5_local_fixture():
    data = load("c5_local.json")
    scoring = o.Scoring(data["alphabet"], tuple(data["table"]), data["opening"], data["extension"])
    for ref, qry, score, cigar, spans in data["pairs"][:24]:
        assert o.scoring_local_align(ref, qry, scoring) == (score, cigar, tuple(spans))

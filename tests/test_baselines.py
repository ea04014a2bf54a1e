"""The CPU baselines the bench times (oracle/baseline.cpp, oracle/baseline_scoring.cpp) against the oracle: a
baseline is a second CPU implementation, so its results must equal the oracle's pair by pair before its time means
anything."""

import ctypes
import os
import random

import numpy as np
import pytest

from dinara_align import synthetic
from oracle import oracle as o

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")


def load(name):
    path = os.path.join(ORACLE_DIR, name)
    if not os.path.exists(path):
        o.build()
    return ctypes.CDLL(path)


def texts(items):
    data = [bytes(x) for x in items]
    ptrs = (ctypes.c_char_p * len(data))(*data)
    lens = (ctypes.c_int64 * len(data))(*[len(x) for x in data])
    return data, ptrs, lens


def test_costs_baseline_matches_oracle():
    lib = load("libbaseline.so")
    i64, vp = ctypes.c_int64, ctypes.c_void_p
    lib.baseline_alignments.argtypes = [i64, vp, vp, vp, vp, i64, i64, i64, i64, vp, vp, i64]
    lib.baseline_alignments.restype = i64
    refs, qrys = synthetic.affine_reads(200, seed=21)
    _, rp, rl = texts(refs)
    _, qp, ql = texts(qrys)
    costs = (ctypes.c_int64 * len(refs))()
    stride = 1200
    cigars = ctypes.create_string_buffer(stride * len(refs))
    lib.baseline_alignments(len(refs), rp, rl, qp, ql, 4, 6, 2, 4, costs, cigars, stride)
    affine = o.Costs.affine(4, 6, 2)
    for i, (r, q) in enumerate(zip(refs, qrys)):
        want = o.costs_align(r, q, affine)
        got = cigars.raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode()
        assert (costs[i], got) == want, i


def test_scoring_baseline_matches_oracle():
    lib = load("libbaseline_scoring.so")
    i64, vp = ctypes.c_int64, ctypes.c_void_p
    lib.baseline_local_alignments.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp, i64, i64, i64, i64, vp, vp, vp, i64]
    lib.baseline_local_alignments.restype = i64
    alphabet, table = synthetic.protein_table()
    scoring = o.Scoring(alphabet, table, -11, -1)
    refs, qrys = synthetic.local_protein(40, seed=31, query=120, window=700)
    rng = random.Random(3)
    # Unrelated and ragged pairs too: ends anywhere, or nothing above zero.
    for _ in range(24):
        refs.append(bytes(rng.choice(alphabet.encode()) for _ in range(rng.randint(1, 90))))
        qrys.append(bytes(rng.choice(alphabet.encode()) for _ in range(rng.randint(1, 90))))
    _, rp, rl = texts(refs)
    _, qp, ql = texts(qrys)
    rcodes = [scoring.codes(r) for r in refs]
    qcodes = [scoring.codes(q) for q in qrys]
    _, rcp, _ = texts(rcodes)
    _, qcp, _ = texts(qcodes)
    size = len(alphabet)
    table32 = np.zeros(32 * 32, dtype=np.int8)
    for a in range(size):
        table32[a * 32:a * 32 + size] = table[a * size:(a + 1) * size]
    scores = (ctypes.c_int64 * len(refs))()
    spans = (ctypes.c_int64 * (4 * len(refs)))()
    stride = 4000
    cigars = ctypes.create_string_buffer(stride * len(refs))
    lib.baseline_local_alignments(len(refs), rp, rl, qp, ql, rcp, qcp, table32.ctypes.data, -12, -1, 1, 4, scores,
                                  spans, cigars, stride)
    for i, (r, q) in enumerate(zip(refs, qrys)):
        want_score, want_cigar, want_spans = o.scoring_local_align(r, q, scoring)
        got = cigars.raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode()
        assert (scores[i], got, tuple(spans[4 * i:4 * i + 4])) == (want_score, want_cigar, want_spans), i


@pytest.mark.parametrize("right", [False, True])
def test_edit_baseline_matches_oracle(right):
    lib = load("libbaseline_edit.so")
    i64, vp = ctypes.c_int64, ctypes.c_void_p
    lib.baseline_edit.argtypes = [i64, vp, vp, vp, vp, i64, i64, vp, vp, i64]
    lib.baseline_edit.restype = i64
    refs, qrys = synthetic.uneven(120, seed=27 + right, clip=(0, 900))
    rng = random.Random(5)
    refs = list(refs) + [b"", b"A", b"", b"ACGT" * 40, bytes(rng.choice(b"ACGT") for _ in range(700))]
    qrys = list(qrys) + [b"", b"", b"C", b"", bytes(rng.choice(b"ACGT") for _ in range(90))]
    _, rp, rl = texts(refs)
    _, qp, ql = texts(qrys)
    costs = (ctypes.c_int64 * len(refs))()
    stride = 4000
    cigars = ctypes.create_string_buffer(stride * len(refs))
    assert lib.baseline_edit(len(refs), rp, rl, qp, ql, 3, int(right), costs, cigars, stride) == len(refs)
    edit = o.Costs.affine(1, 0, 1)
    for i, (r, q) in enumerate(zip(refs, qrys)):
        want = o.costs_align(r, q, edit, right_ties=right)
        got = cigars.raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode()
        assert (costs[i], got) == want, i
    distances = (ctypes.c_int64 * len(refs))()
    lib.baseline_edit(len(refs), rp, rl, qp, ql, 2, 0, distances, None, 0)
    assert list(distances) == [o.edit_distance(r, q) for r, q in zip(refs, qrys)]

"""GPU checks of the learned band proposal: the device featurizer and the bf16 MFMA network against their numpy
restatement (oracle/mlp.py), training on costs the aligner itself computes, and the ablation: with the trained
network's seeds the band certificate proves more pairs in their first sweep, and every distance is unchanged."""

import ctypes
import random

import numpy as np
import pytest

import dinara_align as da
from dinara_align import _lib, synthetic
from oracle import mlp

pytestmark = pytest.mark.gpu

F = mlp.STRIDE


class Batch:
    """A resident distances batch under (4, 6, 2)."""

    def __init__(self, refs, qrys, kind=0, threads=8):
        self.h = _lib.lib()
        self.R, self.Q = _lib.Texts(refs), _lib.Texts(qrys)
        self.costs = _lib.CCosts(4, 6, 2, -1, 0, 0, 0, -1, 0)
        status = ctypes.c_int64(0)
        self.pairs = len(refs)
        self.ptr = self.h.dinara_batch_new(0, kind, self.pairs, *self.R.args(), *self.Q.args(),
                                           ctypes.byref(self.costs), None, None, threads, ctypes.byref(status))
        assert self.ptr, status.value

    def features(self):
        out = np.zeros((self.pairs, F), dtype=np.float32)
        assert self.h.dinara_batch_features(self.ptr, out.ctypes.data) == 0
        return out

    def seed(self, costs):
        arr = np.ascontiguousarray(costs, dtype=np.int64)
        assert self.h.dinara_batch_seed(self.ptr, arr.ctypes.data) == 0

    def distances(self):
        assert self.h.dinara_batch_run(self.ptr) == 0
        out = (ctypes.c_int64 * self.pairs)()
        assert self.h.dinara_batch_distances(self.ptr, out) == 0
        return np.array(out[:], dtype=np.int64), self.h.dinara_batch_stat(self.ptr, 4)

    def free(self):
        self.h.dinara_batch_free(self.ptr)


class Net:
    def __init__(self, seed=1):
        self.h = _lib.lib()
        self.ptr = self.h.dinara_mlp_new(0, seed)
        assert self.ptr
        self.size = self.h.dinara_mlp_parameter_count(self.ptr)

    def weights(self):
        w = np.zeros(self.size, dtype=np.float32)
        assert self.h.dinara_mlp_get_weights(self.ptr, w.ctypes.data) == 0
        return w

    def gradients(self, x, t):
        g = np.zeros(self.size, dtype=np.float32)
        loss = ctypes.c_double(0)
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        assert self.h.dinara_mlp_gradients(self.ptr, len(t), x.ctypes.data, t.ctypes.data, g.ctypes.data,
                                           ctypes.byref(loss)) == 0
        return g, loss.value

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros(len(x), dtype=np.float32)
        assert self.h.dinara_mlp_predict(self.ptr, len(x), x.ctypes.data, out.ctypes.data) == 0
        return out

    def train(self, x, t, epochs=4, batch=256, lr=2e-3):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_double(0)
        assert self.h.dinara_mlp_train(self.ptr, len(t), x.ctypes.data, t.ctypes.data, epochs, batch, lr, 7,
                                       ctypes.byref(loss)) == 0
        return loss.value

    def free(self):
        self.h.dinara_mlp_free(self.ptr)


def random_inputs(rows, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((rows, F), dtype=np.float32)
    x[:, :mlp.FEATURES] = rng.normal(0, 1, (rows, mlp.FEATURES))
    t = rng.normal(6, 1.5, rows).astype(np.float32)
    return x, t


def test_gradients_match_numpy():
    # Rows not a multiple of 32: the padding rows must add nothing.
    net = Net(3)
    try:
        x, t = random_inputs(77, 5)
        got, loss = net.gradients(x, t)
        want, want_loss = mlp.gradients(net.weights(), x, t)
        assert abs(loss - want_loss) <= 1e-4 * max(1.0, abs(want_loss))
        for name, a, b in zip(("w1", "b1", "w2", "b2", "w3", "b3"), mlp.split(got), mlp.split(want)):
            err = np.linalg.norm(a - b)
            assert err <= 1e-3 * max(np.linalg.norm(b), 1e-6) + 1e-6, (name, err, np.linalg.norm(b))
    finally:
        net.free()


def test_forward_matches_numpy():
    net = Net(4)
    try:
        x, _ = random_inputs(200, 6)
        got = net.predict(x)
        want = mlp.forward(net.weights(), x)[0]
        assert np.allclose(got, want, rtol=1e-4, atol=1e-4)
    finally:
        net.free()


@pytest.mark.parametrize("kind", [0, 1])
def test_features_match_numpy(kind):
    # kind 1 (alignments, Ties.LEFT) packs the tape reversed: the features must not change.
    refs, qrys = synthetic.uneven(40, seed=9, clip=(4, 900))
    b = Batch(refs, qrys, kind=kind)
    try:
        got = b.features()
    finally:
        b.free()
    for i, (r, q) in enumerate(zip(refs, qrys)):
        assert np.allclose(got[i], mlp.features(r, q), rtol=1e-5, atol=1e-6), i


def test_training_seeds_prove_more_pairs_and_change_nothing():
    # Costs the aligner itself computes label the training pairs (log(1 + cost)).
    train_r, train_q = synthetic.uneven(6000, seed=21, clip=(100, 1500))
    b = Batch(train_r, train_q)
    x_train = b.features()
    costs, _ = b.distances()
    b.free()
    t_train = np.log1p(costs).astype(np.float32)
    net = Net(5)
    try:
        first = mlp.gradients(net.weights(), x_train[:512], t_train[:512])[1]
        final = net.train(x_train, t_train, epochs=6, batch=256, lr=2e-3)
        assert final < 0.5 * first, (first, final)
        # Fresh pairs: the plain first band against the network's seed, with a margin.
        eval_r, eval_q = synthetic.uneven(3000, seed=22, clip=(100, 1500))
        plain = Batch(eval_r, eval_q)
        plain_costs, plain_first = plain.distances()
        plain.free()
        seeded = Batch(eval_r, eval_q)
        guess = np.expm1(net.predict(seeded.features()) + 0.15)
        seeded.seed(np.maximum(guess, 0).astype(np.int64))
        seeded_costs, seeded_first = seeded.distances()
        seeded.free()
        assert np.array_equal(plain_costs, seeded_costs)
        assert seeded_first > plain_first + 0.3 * len(eval_r), (plain_first, seeded_first)
    finally:
        net.free()


def test_doubling_policy_changes_no_result():
    # Plain band doubling (the baseline the learned band is measured against) proves the same costs and CIGARs.
    refs, qrys = synthetic.uneven(300, seed=23, clip=(50, 2500))
    h = _lib.lib()
    R, Q = _lib.Texts(refs), _lib.Texts(qrys)
    costs = _lib.CCosts(4, 6, 2, -1, 0, 0, 0, -1, 0)
    results = []
    for doubling in (0, 1):
        status = ctypes.c_int64(0)
        b = h.dinara_batch_new(0, 1, len(refs), *R.args(), *Q.args(), ctypes.byref(costs), None, None, 4,
                               ctypes.byref(status))
        assert b, status.value
        assert h.dinara_batch_set_retry(b, doubling) == 0
        assert h.dinara_batch_run(b) == 0
        found = (_lib.CAlignment * len(refs))()
        statuses = (ctypes.c_int64 * len(refs))()
        assert h.dinara_batch_alignments(b, found, statuses) == 0
        results.append([(statuses[i], found[i].cost, ctypes.string_at(found[i].cigar, found[i].cigar_length))
                        for i in range(len(refs))])
        for i in range(len(refs)):
            h.dinara_free(found[i].cigar)
        h.dinara_batch_free(b)
    assert results[0] == results[1]


def test_features_of_long_sequences_match_numpy():
    # Long sequences: a 4-mer counted more than 32,767 times (past any 16-bit count), and ordinary long pairs, in
    # both tape orders.
    rng = random.Random(61)
    long_ref = "".join(rng.choice("ACGT") for _ in range(34000))
    refs = [s.encode() for s in ("A" * 33000, long_ref, "ACGTTGCA" * 40)]
    qrys = [s.encode() for s in ("A" * 10, long_ref[:20000] + "T" * 50 + long_ref[20000:], "ACGTTGCA" * 39)]
    for kind in (0, 1):
        b = Batch(refs, qrys, kind=kind)
        try:
            got = b.features()
        finally:
            b.free()
        for i, (r, q) in enumerate(zip(refs, qrys)):
            assert np.allclose(got[i], mlp.features(r, q), rtol=1e-5, atol=1e-6), (kind, i)


@pytest.mark.parametrize("unit", [False, True])
def test_unbounded_predictions_seed_whole_bands_and_change_nothing(unit):
    # A network whose output is +inf (its last bias), NaN or huge: the guessed cost is clamped before it becomes an
    # integer, the first band is then the widest, and every distance stays the oracle's.
    from oracle import oracle as o
    refs, qrys = synthetic.uneven(200, seed=24, clip=(50, 1200))
    h = _lib.lib()
    R, Q = _lib.Texts(refs), _lib.Texts(qrys)
    costs = _lib.CCosts(1, 0, 1, -1, 0, 0, 0, -1, 0) if unit else _lib.CCosts(4, 6, 2, -1, 0, 0, 0, -1, 0)
    want_costs = o.Costs.affine(1, 0, 1) if unit else o.Costs.affine(4, 6, 2)
    want = [o.costs_cost(r, q, want_costs) for r, q in zip(refs, qrys)]
    net = Net(3)
    try:
        for bias in (np.inf, np.nan, 1e30, -np.inf):
            w = net.weights()
            w[-257:-1] = 0.0  # the output layer's weights: the prediction is its bias alone
            w[-1] = bias
            assert h.dinara_mlp_set_weights(net.ptr, w.ctypes.data) == 0
            status = ctypes.c_int64(0)
            b = h.dinara_batch_new(0, 0, len(refs), *R.args(), *Q.args(), ctypes.byref(costs), None, None, 4,
                                   ctypes.byref(status))
            assert b, status.value
            try:
                assert h.dinara_batch_seed_learned_by_length(b, net.ptr, ctypes.c_double(0.1),
                                                             ctypes.c_double(0.1)) == 0
                assert h.dinara_batch_run(b) == 0
                out = (ctypes.c_int64 * len(refs))()
                assert h.dinara_batch_distances(b, out) == 0
                assert list(out) == want, bias
            finally:
                h.dinara_batch_free(b)
    finally:
        net.free()

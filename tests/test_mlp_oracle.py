"""The learned band proposal's numpy restatement (oracle/mlp.py), pinned before it checks the device: its chain
rule against central differences, its bf16 steps against the unrounded network, its rounding against known values,
and its features against a hand count."""

import numpy as np
import pytest

from oracle import mlp


def small_net(seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(mlp.STRIDE * 256 + 256 + 256 * 256 + 256 + 256 + 1, dtype=np.float32)
    parts = mlp.split(w)
    parts[0][:mlp.FEATURES] = rng.normal(0, np.sqrt(2 / mlp.FEATURES), (mlp.FEATURES, 256))
    parts[1][:] = rng.normal(0, 0.1, 256)
    parts[2][:] = rng.normal(0, np.sqrt(2 / 256), (256, 256))
    parts[3][:] = rng.normal(0, 0.1, 256)
    parts[4][:] = rng.normal(0, np.sqrt(1 / 256), 256)
    parts[5][:] = 0.3
    x = np.zeros((40, mlp.STRIDE), dtype=np.float32)
    x[:, :mlp.FEATURES] = rng.normal(0, 1, (40, mlp.FEATURES))
    t = rng.normal(5, 1, 40).astype(np.float32)
    return w, x, t


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, -2.5, 1 + 2 ** -7 + 2 ** -8, 0.0], dtype=np.float32)
    got = mlp.bf16(v)
    # Halfway cases round to the even neighbour: 1 + 2^-8 to 1, 1 + 2^-7 + 2^-8 to 1 + 2^-6.
    assert list(got) == [1.0, 1.0, 1 + 2 ** -7, -2.5, 1 + 2 ** -6, 0.0]


@pytest.mark.parametrize("seed", [1, 2])
def test_chain_rule_matches_central_differences(seed):
    w, x, t = small_net(seed)
    exact = mlp.exact_gradients(w, x.astype(np.float64), t.astype(np.float64))
    rng = np.random.default_rng(seed + 10)
    h = 256
    # Indices in every tensor: w1 (real inputs), b1, w2, b2, w3, b3.
    starts = [0, mlp.STRIDE * h, mlp.STRIDE * h + h, mlp.STRIDE * h + h + h * h, mlp.STRIDE * h + 2 * h + h * h,
              len(w) - 1]
    picks = [int(rng.integers(0, mlp.FEATURES * h))] + [starts[i] + int(rng.integers(0, 50)) for i in (1, 2, 3, 4)]
    picks.append(starts[5])
    for index in picks:
        numeric = mlp.numeric_gradient(w, x.astype(np.float64), t.astype(np.float64), index)
        assert abs(numeric - exact[index]) <= 1e-5 + 1e-4 * abs(exact[index]), index


def test_bf16_network_follows_the_exact_gradient():
    w, x, t = small_net(3)
    rounded, loss = mlp.gradients(w, x, t)
    exact = mlp.exact_gradients(w, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - mlp.exact_loss(w, x, t)) < 0.05 * mlp.exact_loss(w, x, t)
    # bf16 inputs and activations move each gradient a little, the whole hardly.
    assert np.linalg.norm(rounded - exact) < 0.05 * np.linalg.norm(exact)


def test_features_by_hand():
    f = mlp.features(b"AAAAC", b"AAAA")
    # Reference 4-mers AAAA, AAAC; query AAAA: the difference is one AAAC (code 1).
    scale = 8 / np.sqrt(10)
    assert abs(f[1] - scale) < 1e-6 and f[0] == 0 and np.count_nonzero(f[:256]) == 1
    assert abs(f[256] - (np.log(6) - 7) / 2) < 1e-6 and abs(f[257] - (np.log(5) - 7) / 2) < 1e-6
    assert not f[258:].any()

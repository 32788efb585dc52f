"""CPU checks of the predictor network's oracle and of the predictor's host-side logic.

The oracle (oracle/mlp_oracle.py) is pinned here by mathematics, not by retyping it: its analytic
gradient against central finite differences of its own loss, its forward pass against an explicit
per-sample loop, and Adam against the closed form of a constant-gradient run.  A planted mistake
(the ReLU mask dropped from the backward pass) must fail the finite-difference check.
"""
import os

import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _problem(seed, n=9, i=5, h=7, o=4):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.7, mo.param_count(i, h, o))
    x = rng.normal(0, 1, (n, i))
    t = rng.normal(0, 1, (n, o))
    m = (rng.random((n, o)) < 0.7).astype(np.float64)
    m[0, 0] = 1.0
    return p, x, t, m, (i, h, o)


def _fd_grad(p, x, t, m, dims, eps=1e-6):
    g = np.zeros_like(p)
    for j in range(p.size):
        a, b = p.copy(), p.copy()
        a[j] += eps
        b[j] -= eps
        g[j] = (mo.loss(a, x, t, m, *dims) - mo.loss(b, x, t, m, *dims)) / (2 * eps)
    return g


def _far_from_kinks(p, x, dims, margin=1e-3):
    w1, b1, _, _ = mo.unpack(p, *dims)
    return np.abs(x @ w1.T + b1).min() > margin


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_gradient_matches_finite_differences(seed):
    p, x, t, m, dims = _problem(seed)
    assert _far_from_kinks(p, x, dims)
    _, _, g = mo.backward(p, x, t, m, *dims)
    fd = _fd_grad(p, x, t, m, dims)
    assert np.abs(g - fd).max() <= 1e-6 * max(1.0, np.abs(fd).max())
    assert np.abs(g).max() > 1e-3  # non-trivial


def test_planted_backward_mistake_fails_finite_differences():
    p, x, t, m, dims = _problem(3)
    w1, b1, w2, b2 = mo.unpack(p, *dims)
    hid, y = mo.forward(p, x, *dims)
    dy = 2.0 * (y - t) * m / m.sum()
    dh_wrong = dy @ w2  # ReLU mask dropped
    g_wrong = np.concatenate([(dh_wrong.T @ x).ravel(), dh_wrong.sum(0), (dy.T @ hid).ravel(),
                              dy.sum(0)])
    fd = _fd_grad(p, x, t, m, dims)
    assert np.abs(g_wrong - fd).max() > 1e-3


def test_oracle_forward_matches_per_sample_loop():
    p, x, _, _, (i, h, o) = _problem(4)
    w1, b1, w2, b2 = mo.unpack(p, i, h, o)
    hid, y = mo.forward(p, x, i, h, o)
    for s in range(len(x)):
        a = [sum(w1[u, k] * x[s, k] for k in range(i)) + b1[u] for u in range(h)]
        r = [max(v, 0.0) for v in a]
        assert np.allclose(hid[s], r, rtol=0, atol=1e-12)
        yy = [sum(w2[q, u] * r[u] for u in range(h)) + b2[q] for q in range(o)]
        assert np.allclose(y[s], yy, rtol=0, atol=1e-12)


def test_adam_constant_gradient_closed_form():
    # with a constant gradient the bias-corrected moments are g and g^2, so with eps = 0 every
    # step moves each parameter by exactly lr * sign(g)
    rng = np.random.default_rng(5)
    p0 = rng.normal(0, 1, 11)
    g = rng.normal(0, 1, 11)
    p, m, v = p0.copy(), np.zeros(11), np.zeros(11)
    for step in range(1, 6):
        p, m, v = mo.adam(p, g, m, v, step, lr=0.01, eps=0.0)
        assert np.allclose(p, p0 - 0.01 * step * np.sign(g), rtol=0, atol=1e-12)


def test_adam_without_bias_correction_would_differ():
    g = np.array([0.5])
    p, m, v = mo.adam(np.zeros(1), g, np.zeros(1), np.zeros(1), 1, lr=0.1, eps=0.0)
    raw = -0.1 * (0.1 * 0.5) / np.sqrt(0.001 * 0.25)  # uncorrected moments
    assert abs(p[0] - (-0.1)) < 1e-12 and abs(raw - (-0.1)) > 1e-3


def test_oracle_training_decreases_loss():
    p, x, t, m, dims = _problem(6, n=32)
    _, losses = mo.train(p, x, t, m, *dims, steps=200, lr=1e-2)
    assert losses[-1] < 0.5 * losses[0]


# ---- predictor host logic ---------------------------------------------------------------------

def test_features_closed_forms():
    from huffman_coding_gap_arrays_amd import predictor as pr

    f = pr.features(np.full(256, 10), np.full(256, 8))
    assert np.allclose(f, [8.0, np.log2(2560), 8.0, 8.0, 8.0])
    h = np.zeros(256)
    h[7] = 100
    ln = np.zeros(256)
    ln[7] = 1
    assert np.allclose(pr.features(h, ln), [0.0, np.log2(100), 1.0, 1.0, 0.0])
    assert np.allclose(pr.features(np.zeros(256), np.zeros(256)), 0.0)
    h = np.zeros(256)
    h[:4] = [4, 2, 1, 1]
    ln = np.zeros(256)
    ln[:4] = [1, 2, 3, 3]
    assert np.allclose(pr.features(h, ln), [1.75, 3.0, 3.0, 1.75, 2.0])


def test_targets_and_regret():
    from huffman_coding_gap_arrays_amd import predictor as pr

    enc = np.full((3, pr.N_ENC), 2.0)
    enc[:, 5] = 1.0
    enc[1, 7] = np.nan
    enc[2, 9] = -1.0
    dec = np.full((3, pr.N_DEC), 4.0)
    dec[:, 0] = 1.0
    T, M = pr.targets(enc, dec)
    assert T.shape == (3, pr.N_OUT) and M.shape == (3, pr.N_OUT)
    assert (T[:, 5] == 0).all() and np.allclose(T[0, 0], 1.0)
    assert T[1, 7] == pr.UNSUPPORTED_PENALTY and M[1, 7] == 1.0
    assert M[2, 9] == 0.0 and M.sum() == 3 * pr.N_OUT - 1
    assert np.allclose(T[:, pr.N_ENC + 1], 2.0)
    r = pr.regret(enc[:2], [5, 5])
    assert r["aggregate"] == 0.0 and r["mean"] == 0.0
    r = pr.regret(enc[:2], [0, 5])
    assert np.isclose(r["aggregate"], 0.5) and np.isclose(r["mean"], 0.5)
    with pytest.raises(ValueError):
        pr.regret(enc[:2], [5, 7])


def test_candidates_and_defaults():
    from huffman_coding_gap_arrays_amd import predictor as pr

    assert pr.N_ENC == 27 and pr.N_DEC == 45
    assert pr.ENC_CANDIDATES[pr.candidate_index(pr.DEFAULT_ENC, pr.ENC_CANDIDATES)] == \
        pr.DEFAULT_ENC
    assert pr.DEC_CANDIDATES[pr.candidate_index(pr.DEFAULT_DEC, pr.DEC_CANDIDATES)] == \
        pr.DEFAULT_DEC


def test_default_shapes_match_the_library(capi):
    from huffman_coding_gap_arrays_amd import _capi
    from huffman_coding_gap_arrays_amd import predictor as pr

    d = _capi.default_config().as_dict()
    assert d["prefix_bits"] == 0  # 0 = the table builder's default first-level width, 11 bits
    d["prefix_bits"] = 11
    for k, v in {**pr.DEFAULT_ENC, **pr.DEFAULT_DEC}.items():
        assert d[k] == v, k


def test_sweep_spec_is_deterministic_and_in_range():
    from huffman_coding_gap_arrays_amd import predictor as pr

    a, b = pr.sweep_spec(64, 11), pr.sweep_spec(64, 11)
    assert a == b
    assert all(4096 <= d["n"] <= 256 << 20 for d in a)
    assert {d["kind"] for d in pr.sweep_spec(512, 11)} == {"zipf", "geometric", "spike"}


def test_split_is_disjoint_and_complete():
    from huffman_coding_gap_arrays_amd import predictor as pr

    sw = dict(X=np.zeros((21, 5)))
    tr, he = pr.split(sw)
    assert len(set(tr) & set(he)) == 0 and sorted(list(tr) + list(he)) == list(range(21))
    assert len(he) == 5


@pytest.mark.skipif(not os.path.exists(os.path.join(
    os.path.dirname(__file__), "..", "huffman_coding_gap_arrays_amd", "data",
    "launch_sweeps.npz")), reason="no committed sweep data")
def test_committed_sweeps_are_well_formed():
    from huffman_coding_gap_arrays_amd import predictor as pr

    sw = pr.load_sweeps()
    N = len(sw["X"])
    assert N >= 256
    assert sw["enc_ms"].shape == (N, pr.N_ENC) and sw["dec_ms"].shape == (N, pr.N_DEC)
    # every buffer has at least one runnable shape per stage, and the default shapes always run
    assert (~np.isnan(sw["enc_ms"])).any(1).all() and (~np.isnan(sw["dec_ms"])).any(1).all()
    de = pr.candidate_index(pr.DEFAULT_ENC, pr.ENC_CANDIDATES)
    dd = pr.candidate_index(pr.DEFAULT_DEC, pr.DEC_CANDIDATES)
    assert not np.isnan(sw["enc_ms"][:, de]).any() and not np.isnan(sw["dec_ms"][:, dd]).any()
    # features recomputed from the spec's sizes agree
    assert np.allclose(sw["X"][:, 1], np.log2(sw["n"]))
    spec = pr.sweep_spec(N, 11)
    assert [d["n"] for d in spec] == list(sw["n"])

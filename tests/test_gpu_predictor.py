"""The predictor network's device kernels against the CPU oracle, and what the trained predictor
gains over the default launch shapes.  Needs an MI355X.

Tolerances (fp32 device kernels vs the float64 oracle):
  forward:   |y_gpu - y_ref| <= 1e-5 * (1 + |y_ref|)
  gradient:  |g_gpu - g_ref| <= 1e-4 * max|g_ref|      (sums over <= 512 samples)
  Adam:      |p_gpu - p_ref| <= 1e-6 * (1 + |p_ref|)
"""
import os
import sys

import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SWEEPS = os.path.join(HERE, "..", "huffman_coding_gap_arrays_amd", "data", "launch_sweeps.npz")
MODEL = os.path.join(HERE, "..", "huffman_coding_gap_arrays_amd", "data", "predictor.npz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem(seed, n, i, h, o):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.5, mo.param_count(i, h, o))
    x = rng.normal(0, 1, (n, i))
    t = rng.normal(0, 1, (n, o))
    m = (rng.random((n, o)) < 0.8).astype(np.float64)
    m[0, 0] = 1.0
    return p, x, t, m


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _host(t):
    return t.cpu().numpy().astype(np.float64)


SHAPES = [(1, 1, 1, 1), (7, 5, 3, 4), (64, 5, 32, 72), (512, 5, 32, 72), (300, 9, 17, 5)]


@pytest.mark.parametrize("n,i,h,o", SHAPES)
def test_forward_matches_oracle(cuda, n, i, h, o):
    from huffman_coding_gap_arrays_amd.predictor import Network

    say("forward", n, i, h, o)
    p, x, _, _ = _problem(n + i + h + o, n, i, h, o)
    p32 = p.astype(np.float32).astype(np.float64)
    x32 = x.astype(np.float32).astype(np.float64)
    net = Network(i, h, o, params=p)
    hid, y = net.forward(_dev(x))
    rh, ry = mo.forward(p32, x32, i, h, o)
    assert np.all(np.abs(_host(y) - ry) <= 1e-5 * (1 + np.abs(ry)))
    assert np.all(np.abs(_host(hid) - rh) <= 1e-5 * (1 + np.abs(rh)))


@pytest.mark.parametrize("n,i,h,o", SHAPES)
def test_gradient_matches_oracle(cuda, n, i, h, o):
    from huffman_coding_gap_arrays_amd.predictor import Network

    say("gradient", n, i, h, o)
    p, x, t, m = _problem(7 * n + i, n, i, h, o)
    f = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    net = Network(i, h, o, params=p)
    net.gradient(_dev(x), _dev(t), _dev(m))
    _, _, g = mo.backward(f(p), f(x), f(t), f(m), i, h, o)
    gg = _host(net.grad)
    assert np.abs(gg - g).max() <= 1e-4 * max(np.abs(g).max(), 1e-30)


def test_adam_step_matches_oracle(cuda):
    import torch

    from huffman_coding_gap_arrays_amd.predictor import Network

    say("adam")
    i, h, o = 5, 32, 72
    p, x, t, m = _problem(11, 64, i, h, o)
    net = Network(i, h, o, params=p)
    rng = np.random.default_rng(3)
    pr = net.host_params().astype(np.float64)
    mr, vr = np.zeros_like(pr), np.zeros_like(pr)
    for step in range(1, 4):
        g = rng.normal(0, 1e-2, pr.size).astype(np.float32)
        net.grad.copy_(torch.from_numpy(g))
        net.adam(1e-3)
        pr, mr, vr = mo.adam(pr, g.astype(np.float64), mr, vr, step, 1e-3)
        assert np.all(np.abs(_host(net.params) - pr) <= 1e-6 * (1 + np.abs(pr)))


def test_device_training_tracks_the_oracle(cuda):
    from huffman_coding_gap_arrays_amd.predictor import Network

    say("training")
    n, i, h, o = 128, 5, 16, 6
    rng = np.random.default_rng(21)
    x = rng.normal(0, 1, (n, i))
    w = rng.normal(0, 1, (o, i))
    t = np.tanh(x @ w.T)  # a smooth target the network can fit
    m = np.ones((n, o))
    net = Network(i, h, o, seed=4)
    p0 = net.host_params().astype(np.float64)
    hist = net.fit(_dev(x), _dev(t), _dev(m), steps=300, lr=1e-2, log_every=50)
    _, ref = mo.train(p0, x.astype(np.float32).astype(np.float64),
                      t.astype(np.float32).astype(np.float64), m, i, h, o, steps=300, lr=1e-2)
    for s, loss in hist:
        assert abs(loss - ref[s]) <= 0.05 * ref[s] + 1e-4, (s, loss, ref[s])
    assert hist[-1][1] < 0.2 * hist[0][1]


@pytest.fixture(scope="module")
def sweeps():
    if not os.path.exists(SWEEPS):
        pytest.skip("no committed sweep data")
    from huffman_coding_gap_arrays_amd import predictor as pr

    return pr.load_sweeps(SWEEPS)


def test_trained_predictor_beats_the_default_shapes(cuda, sweeps):
    """Trained on the GPU from the training rows of the committed sweeps, judged on the held-out
    rows by the exhaustive tuner's own timings: the learned shapes cost less total time than the
    default shapes, encode and decode together."""
    from huffman_coding_gap_arrays_amd import predictor as pr

    say("train on sweeps")
    tr, held = pr.split(sweeps)
    pred, loss = pr.Predictor.train(sweeps["X"][tr], sweeps["enc_ms"][tr],
                                    sweeps["dec_ms"][tr], steps=3000, lr=3e-3, log_every=500)
    assert loss[-1][1] < 0.5 * loss[0][1]
    ev = pr.evaluate(pred, sweeps, held)
    say(ev)
    learned = ev["encode"]["learned"]["chosen_ms"] + ev["decode"]["learned"]["chosen_ms"]
    default = ev["encode"]["default"]["chosen_ms"] + ev["decode"]["default"]["chosen_ms"]
    best = ev["encode"]["learned"]["best_ms"] + ev["decode"]["learned"]["best_ms"]
    assert learned < default
    # closer to the exhaustive optimum than the default is
    assert learned / best - 1 < 0.5 * (default / best - 1)


def test_shipped_model_chooses_supported_shapes_and_batch_is_exact(cuda):
    if not os.path.exists(MODEL):
        pytest.skip("no shipped model")
    import torch

    from huffman_coding_gap_arrays_amd import batch, predictor as pr

    say("shipped model")
    model = pr.Predictor.load(MODEL)
    arena, bufs, spec = batch.make_workload(count=48, seed=9, max_bytes=4 << 20)
    hists = batch.histograms(bufs)
    cbs = batch.codebooks(hists)
    enc, dec, learned = pr.choose_configs(hists, cbs, model, buffers=bufs, fallback=True)
    assert len(enc) == len(dec) == len(bufs)
    bc = batch.BatchCodec(bufs, configs=enc, dec_configs=dec, hists=hists)
    bc.run()
    torch.cuda.synchronize()
    assert bc.mismatches() == 0
    bp = batch.BatchCodec(bufs, predictor=model)
    bp.run()
    torch.cuda.synchronize()
    assert bp.mismatches() == 0
    for i in range(len(bufs)):
        assert np.array_equal(bc.payload_u32(i), bp.payload_u32(i))

"""Host-side logic of the serving loop: memory bounds, arrivals, workload, statistics."""
import math

import numpy as np
import pytest


def test_encode_reservation_bound_covers_exact_sizes(capi):
    import ctypes

    from huffman_coding_gap_arrays_amd import _capi, serve

    lib = _capi.load()
    rng = np.random.default_rng(3)
    for trial in range(200):
        n = int(rng.integers(1, 1 << 24))
        k = int(rng.integers(1, 257))
        hist = np.zeros(256, np.uint32)
        w = rng.pareto(0.7, k) + 1e-3
        c = np.maximum(1, np.floor(w / w.sum() * n)).astype(np.uint64)
        c[0] += max(0, n - int(c.sum()))
        hist[rng.choice(256, k, replace=False)] = c.astype(np.uint32)
        cb = _capi.Codebook()
        assert lib.hgap_codebook_build(hist.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cb)) == 0
        sz = _capi.Sizes()
        assert lib.hgap_encode_sizes(ctypes.byref(cb), hist.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.byref(sz)) == 0
        total = int(hist.sum())
        assert serve.encode_reserve_exact(total, sz) <= serve.encode_reserve_bound(total)
    # longest codes (16 bits for every byte) meet the bound's payload term exactly
    n = 1 << 20
    sz = _capi.Sizes()
    sz.payload_words = (16 * n + 31) // 32
    sz.gap_words = ((16 * n + 127) // 128 + 7) // 8
    assert serve.encode_reserve_exact(n, sz) == serve.encode_reserve_bound(n)


def test_poisson_arrivals():
    from huffman_coding_gap_arrays_amd import serve

    a = serve.poisson_arrivals(20000, 1000.0, seed=3)
    assert a == serve.poisson_arrivals(20000, 1000.0, seed=3)
    assert all(y > x for x, y in zip(a, a[1:]))
    gaps = np.diff([0.0] + a)
    # exponential inter-arrival times: mean 1/rate, coefficient of variation 1
    assert abs(gaps.mean() - 1e-3) < 0.03e-3
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.03
    assert serve.poisson_arrivals(5, math.inf) == [0.0] * 5


def test_workload_spec():
    from huffman_coding_gap_arrays_amd import serve

    spec = serve.workload_spec(4000, seed=5)
    assert spec == serve.workload_spec(4000, seed=5)
    assert all(1024 <= d["n"] <= 1 << 28 for d in spec)
    enc = sum(d["kind"] == "encode" for d in spec)
    assert 0.45 < enc / len(spec) < 0.55
    assert all(0.6 <= d["s"] <= 2.0 for d in spec)
    # log-uniform sizes: mean (2^28 - 2^10) / ln(2^18) ~ 21.5 MB
    mean = np.mean([d["n"] for d in spec])
    assert 15e6 < mean < 30e6


def test_percentile_and_stats():
    from huffman_coding_gap_arrays_amd import serve

    xs = list(range(1, 101))
    assert serve.percentile(xs, 50) == 50 and serve.percentile(xs, 99) == 99
    assert serve.percentile(xs, 100) == 100 and serve.percentile([7], 99) == 7
    reqs = [serve.Request(i, "encode" if i % 2 else "decode", 1000, arrival=i * 1e-3,
                          done=i * 1e-3 + 2e-3) for i in range(10)]
    reqs.append(serve.Request(10, "encode", 5, arrival=0.0, done=0.0, error="too big"))
    st = serve.stats(reqs)
    assert st["served"] == 10 and st["rejected"] == 1
    assert abs(st["p50_ms"] - 2.0) < 1e-9 and abs(st["p99_ms"] - 2.0) < 1e-9
    assert st["bytes"] == 10000 and abs(st["makespan_s"] - 0.011) < 1e-9
    assert st["encode"]["count"] == 5 and st["decode"]["count"] == 5


def test_implied_counts_follow_code_lengths(capi):
    from huffman_coding_gap_arrays_amd import codec, serve

    hist = np.zeros(256, np.uint32)
    hist[[3, 9, 11, 200]] = [500, 250, 125, 125]
    cb = codec.codebook_build(hist)
    sz = codec.encode_sizes(cb, hist)
    e = codec.Encoded(1000, cb, sz, None, None)
    h = serve.implied_counts(e)
    assert np.allclose(h[[3, 9, 11, 200]], [500, 250, 125, 125])
    assert h.sum() == 1000

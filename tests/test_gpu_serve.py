"""Serving loop on the device: every response is the same alone and in the mix.  Needs an
MI355X.  Tolerance: bit-exact (encode payload and gap words, decode bytes)."""
import sys

import numpy as np
import pytest
import torch

from huffman_coding_gap_arrays_amd import serve

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _snapshot(reqs):
    out = {}
    for r in reqs:
        assert r.error is None, (r.rid, r.error)
        if r.kind == "encode":
            e = r.result
            out[r.rid] = (e.payload_u32().copy(), e.gap_u32().copy(),
                          tuple(e.codebook.length[:]))
        else:
            out[r.rid] = r.result[:r.n].cpu().numpy().copy()
    return out


@pytest.fixture(scope="module")
def reqs(cuda):
    spec = serve.workload_spec(44, seed=21, min_bytes=1024, max_bytes=3 << 20)
    for i, n in enumerate((1, 15, 16, 17, 4095, 65537)):
        spec.append(dict(rid=44 + i, kind="encode" if i % 2 else "decode", n=n, s=1.3,
                         seed=900 + i, perm_seed=77 + i))
    say("making", len(spec), "requests")
    return serve.make_requests(spec)


@pytest.fixture(scope="module")
def alone(reqs):
    say("alone")
    srv = serve.Server(cap_bytes=1 << 34, max_batch=1, max_inflight=1, streams=1)
    serve.reset(reqs, [0.0] * len(reqs))
    st = srv.run(reqs)
    assert st["served"] == len(reqs) and st["mean_batch"] == 1.0
    assert serve.verify(reqs) == []
    return _snapshot(reqs)


def test_encode_responses_match_oracle(reqs, alone, orc):
    say("oracle")
    for r in reqs:
        if r.kind != "encode":
            continue
        h = r.data.cpu().numpy()
        cb = orc.codebook(orc.histogram(h))
        pay, gap = orc.encode(h, cb)
        got_pay, got_gap, _ = alone[r.rid]
        assert np.array_equal(got_pay, pay), r.rid
        assert np.array_equal(got_gap, gap), r.rid


@pytest.mark.parametrize("streams,inflight,batch,chained_min",
                         [(2, 3, 8, 512), (3, 4, 64, 0), (1, 2, 4, 512)])
def test_same_alone_and_in_the_mix(reqs, alone, streams, inflight, batch, chained_min):
    """chained_min = 0: every encode launch takes the chained encoder (the requests alone take
    the per-tile look-back): the same bits."""
    needs = [serve.request_bytes(r) for r in reqs]
    cap = int(2.5 * max(needs))  # tight: admission has to wait for memory
    say("mix", streams, inflight, batch, chained_min, "cap", cap)
    srv = serve.Server(cap_bytes=cap, max_batch=batch, max_inflight=inflight, streams=streams,
                       chained_min_ranges=chained_min)
    serve.reset(reqs, serve.poisson_arrivals(len(reqs), 200000.0, seed=streams))
    st = srv.run(reqs)
    assert st["served"] == len(reqs) and st["rejected"] == 0
    assert srv.peak_reserved <= cap and srv.reserved == 0
    assert st["mean_batch"] > 1.0 and srv.peak_inflight > 1
    assert st["p50_ms"] <= st["p99_ms"]
    got = _snapshot(reqs)
    for r in reqs:
        if r.kind == "encode":
            a, b = alone[r.rid], got[r.rid]
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), r.rid
            assert a[2] == b[2], r.rid
        else:
            assert np.array_equal(alone[r.rid], got[r.rid]), r.rid
    assert serve.verify(reqs) == []


def test_request_over_the_cap_is_rejected(reqs):
    say("reject")
    small = [r for r in reqs if r.n <= 4096]
    big = max(reqs, key=lambda r: r.n)
    sel = small + [big]
    cap = max(serve.request_bytes(r) for r in small) * 4
    assert serve.request_bytes(big) > cap
    srv = serve.Server(cap_bytes=cap, max_batch=4, max_inflight=2, streams=2)
    serve.reset(sel, [0.0] * len(sel))
    st = srv.run(sel)
    assert st["rejected"] == 1 and big.error is not None and big.result is None
    assert st["served"] == len(small)
    assert serve.verify(small) == []


def test_decode_tables_stay_on_the_device(reqs):
    say("table cache")
    dec = [r for r in reqs if r.kind == "decode"]
    twice = dec + [serve.Request(1000 + r.rid, "decode", r.n, data=r.data,
                                 compressed=r.compressed) for r in dec]
    srv = serve.Server(max_batch=4, max_inflight=2, streams=2)
    serve.reset(twice, [0.0] * len(twice))
    srv.run(twice)
    assert srv.tables.misses == len({srv.tables.key(r.compressed.codebook) for r in dec})
    assert srv.tables.hits >= len(dec)
    assert serve.verify(twice) == []


def test_mix_with_learned_shapes(reqs, alone):
    import os

    from huffman_coding_gap_arrays_amd import predictor

    if not os.path.exists(predictor.DEFAULT_MODEL):
        pytest.skip("no shipped model")
    say("learned shapes")
    srv = serve.Server(max_batch=16, max_inflight=3, streams=2,
                       predictor=predictor.Predictor.load())
    serve.reset(reqs, serve.poisson_arrivals(len(reqs), 20000.0, seed=9))
    st = srv.run(reqs)
    assert st["served"] == len(reqs)
    got = _snapshot(reqs)
    for r in reqs:
        if r.kind == "encode":
            assert np.array_equal(alone[r.rid][0], got[r.rid][0]), r.rid
            assert np.array_equal(alone[r.rid][1], got[r.rid][1]), r.rid
        else:
            assert np.array_equal(alone[r.rid], got[r.rid]), r.rid
    torch.cuda.synchronize()

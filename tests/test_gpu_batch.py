"""Batched encode/decode (many buffers, own codebooks, one launch per shape group, HIP graph
replay) against the oracle and the per-buffer path.  Needs an MI355X."""
import sys

import numpy as np
import pytest

from huffman_coding_gap_arrays_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def batch_data(cuda):
    import torch

    from huffman_coding_gap_arrays_amd import batch, codec

    spec = synth.batch_spec(count=60, seed=9, min_bytes=1 << 10, max_bytes=3 << 20)
    spec += [dict(n=n, s=1.2, seed=1000 + n, perm_seed=n) for n in (1, 15, 16, 17, 4095)]
    spec.append(dict(n=20000, s=40.0, seed=77, perm_seed=1))  # one distinct byte
    # empty buffers (no bytes, no ranges, no tiles) between and after the others
    spec.insert(30, dict(n=0, s=1.2, seed=3, perm_seed=3))
    spec.append(dict(n=0, s=1.2, seed=4, perm_seed=4))
    arena = batch.Arena([d["n"] for d in spec])
    bufs = []
    for i, d in enumerate(spec):
        b = arena.view(i)
        codec.synth_fill(b, d["seed"], synth.cdf_from_pmf(synth.pmf_zipf(d["s"])),
                         synth.permutation(d["perm_seed"]))
        bufs.append(b)
    torch.cuda.synchronize()
    return batch, codec, arena, bufs


def test_batched_histograms(batch_data):
    batch, codec, arena, bufs = batch_data
    h = batch.histograms(bufs)
    for i, b in enumerate(bufs):
        assert (h[i] == np.bincount(b.cpu().numpy(), minlength=256)).all(), i


def test_batched_range_histograms(batch_data):
    """Every buffer's histogram and the counts of each of its 256 KiB ranges (ragged last
    range, one-range buffers) equal numpy's."""
    batch, codec, arena, bufs = batch_data
    h, rh, base = batch.histograms_ranges(bufs)
    rows = rh.cpu().numpy().view(np.uint32)
    R = codec.RANGE_BYTES
    for i, b in enumerate(bufs):
        d = b.cpu().numpy()
        assert (h[i] == np.bincount(d, minlength=256)).all(), i
        for r in range(-(-d.size // R)):
            want = np.bincount(d[r * R:(r + 1) * R], minlength=256)
            assert (rows[base[i] + r] == want).all(), (i, r)


def test_batched_codebooks_match_single(batch_data, orc):
    batch, codec, arena, bufs = batch_data
    h = batch.histograms(bufs)
    cbs = batch.codebooks(h, threads=4)
    for i in range(len(bufs)):
        o = orc.codebook(h[i].astype(np.uint64))
        assert list(cbs[i].code) == list(o["code"]) and list(cbs[i].length) == list(o["length"])


@pytest.mark.parametrize("chained", [False, True])
@pytest.mark.parametrize("mixed", [False, True])
def test_batched_encode_decode_bit_exact(batch_data, orc, mixed, chained):
    """Per-tile look-back (chained=False) and chained encode launches (chained=True: the range
    counts of the batched histogram give every range's offset) write the oracle's streams."""
    import torch

    from huffman_coding_gap_arrays_amd import tuner

    batch, codec, arena, bufs = batch_data
    configs = None
    if mixed:  # every buffer its own launch shape
        space_e = list(tuner.enc_space())
        space_d = list(tuner.dec_space())
        configs = [{**space_e[(7 * i) % len(space_e)], **{k: v for k, v in
                   space_d[(5 * i) % len(space_d)].items() if k != "lookback_window"}}
                   for i in range(len(bufs))]
        # LDS limits: drop shapes that cannot hold 16-bit codes with 1024 x 64 B tiles
        for c in configs:
            if c["enc_threads"] * c["enc_bytes_per_thread"] > 32768:
                c["enc_bytes_per_thread"] = 32
    say("batch codec", "mixed" if mixed else "default", "chained" if chained else "look-back")
    bc = batch.BatchCodec(bufs, configs=configs, chained=chained)
    assert all(L["chained"] == chained for L in bc.enc_launches)
    bc.run()
    torch.cuda.synchronize()
    for i, b in enumerate(bufs):
        d = b.cpu().numpy()
        if d.size == 0:
            continue
        ocb = orc.codebook(orc.histogram(d))
        pay, gap = orc.encode(d, ocb)
        assert (bc.payload_u32(i) == pay).all(), (i, d.size)
        assert (bc.gap_u32(i) == gap).all(), (i, d.size)
    assert bc.mismatches() == 0


def test_batched_decode_table_kinds_and_density_classes(batch_data):
    """The batch mixes one-level and two-level decode tables (the kernels pick the walk per
    tile) and spans several density classes (one decode launch each); all decode exactly."""
    import torch

    batch, codec, arena, bufs = batch_data
    bc = batch.BatchCodec(bufs)
    live = bc.ns > 0
    l2 = bc.dec_info["l2_entries"][live]
    assert (l2 == 0).any() and (l2 > 0).any()
    expect = batch.expected_symbols(bc.ns, bc.gap_elements,
                                    bc.dec_info["max_symbols_per_segment"])
    classes = set(batch.density_class(expect[live]).tolist())
    assert len(classes) >= 2
    assert len(bc.dec_launches) == len(classes)
    bc.run()
    torch.cuda.synchronize()
    assert bc.mismatches() == 0


def test_batched_decode_through_small_image(batch_data, monkeypatch):
    """A batched decode whose LDS image is sized for one codeword per segment sends every tile
    through the image in windows; every buffer still decodes to its input."""
    import torch

    batch, codec, arena, bufs = batch_data
    monkeypatch.setattr(batch, "expected_symbols",
                        lambda n, g, m: np.ones(np.asarray(n).shape, np.int64))
    say("batch codec, one-codeword image")
    bc = batch.BatchCodec(bufs)
    bc.run()
    torch.cuda.synchronize()
    assert bc.mismatches() == 0


@pytest.mark.parametrize("chained", [False, True])
def test_batched_graph_replays(batch_data, chained):
    """The captured graph re-runs encode + decode exactly, replay after replay, with eager work
    and synchronizations in between."""
    import torch

    batch, codec, arena, bufs = batch_data
    bc = batch.BatchCodec(bufs, chained=chained)
    bc.run()
    torch.cuda.synchronize()
    ref_pay = bc.payload.mem.clone()
    ref_gap = bc.gap.mem.clone()
    g = bc.capture()
    for r in range(3):
        bc.out.mem.zero_()
        bc.payload.mem.fill_(0xA5)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert bc.mismatches() == 0, r
        for i in range(len(bufs)):
            o, s = bc.payload.offsets[i], bc.payload.sizes[i]
            assert torch.equal(bc.payload.mem[o:o + s], ref_pay[o:o + s]), (r, i)
            o, s = bc.gap.offsets[i], bc.gap.sizes[i]
            assert torch.equal(bc.gap.mem[o:o + s], ref_gap[o:o + s]), (r, i)
        _ = (bufs[0].float().sum() + 1).item()  # eager work between replays


def test_plain_loop_matches(batch_data):
    import torch

    batch, codec, arena, bufs = batch_data
    outs = batch.plain_loop(bufs[:20])
    for o, b in zip(outs, bufs[:20]):
        assert torch.equal(o, b)

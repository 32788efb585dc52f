"""HIP kernels (through the C ABI) against the oracle, bit-exact.  Needs an MI355X."""
import ctypes
import os
import sys

import numpy as np
import pytest

from huffman_coding_gap_arrays_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def codec(cuda):
    from huffman_coding_gap_arrays_amd import codec as c

    return c


def device_data(codec, n, seed, cdf, perm, start=0):
    import torch

    d = torch.empty(max(n, 16), dtype=torch.uint8, device="cuda")
    if n:
        codec.synth_fill(d[:n], seed, cdf, perm, start=start)
    return d[:n]


def dist_spec(dist):
    if dist == "zipf":
        return synth.cdf_from_pmf(synth.pmf_zipf(1.1)), synth.permutation(77)
    if dist == "geometric":
        return synth.cdf_from_pmf(synth.pmf_geometric(0.1)), synth.IDENTITY
    if dist == "skew":
        return synth.cdf_from_pmf(synth.pmf_skew(0.9)), synth.IDENTITY
    if dist == "uniform":
        return synth.cdf_from_pmf(synth.pmf_uniform()), synth.IDENTITY
    if dist == "gauss":
        return synth.cdf_from_pmf(synth.pmf_gauss_q8()), synth.GAUSS_BYTES
    if dist == "k1":
        p = np.zeros(256)
        p[42] = 1
        return synth.cdf_from_pmf(p), synth.IDENTITY
    if dist == "k2":
        p = np.zeros(256)
        p[[7, 200]] = [0.7, 0.3]
        return synth.cdf_from_pmf(p), synth.IDENTITY
    if dist == "fib":  # Fibonacci-like probabilities: 16-bit codes, two-level decode table
        fib = [1.0, 1.0]
        while len(fib) < 40:
            fib.append(fib[-1] + fib[-2])
        p = np.zeros(256)
        p[:40] = fib[::-1]
        return synth.cdf_from_pmf(p), synth.permutation(5)
    raise ValueError(dist)


def test_synth_fill_matches_oracle(codec, orc):
    for name in synth.CONFIGS:
        n, seed, cdf, perm = synth.spec(name)
        for start in (0, 999_999_937):
            say("synth", name, start)
            d = device_data(codec, 100_003, seed, cdf, perm, start).cpu().numpy()
            assert (d == orc.synth(100_003, start, seed, cdf, perm)).all(), name


@pytest.mark.parametrize("n", [0, 1, 15, 16, 17, 4095, 65537, (1 << 20) + 3, 16 << 20])
def test_histogram(codec, n):
    say("histogram n =", n)
    cdf, perm = dist_spec("zipf")
    d = device_data(codec, n, 3, cdf, perm)
    h = codec.histogram(d)
    want = np.bincount(d.cpu().numpy(), minlength=256)
    assert (h.cpu().numpy().astype(np.int64) == want).all()


RANGE = 256 * 1024  # HGAP_RANGE_BYTES
RANGE_SIZES = [1, 15, 16, 17, 4095, 65536, RANGE - 1, RANGE, RANGE + 1, 3 * RANGE - 17,
               (5 << 20) + 123]


def range_counts(d, n):
    rows = -(-n // RANGE)
    want = np.zeros((rows, 256), dtype=np.int64)
    for r in range(rows):
        want[r] = np.bincount(d[r * RANGE: (r + 1) * RANGE], minlength=256)
    return want


@pytest.mark.parametrize("dist", ["zipf", "skew", "fib", "k1"])
def test_histogram_ranges_and_chained_encode(codec, orc, dist):
    """Per-range counts equal numpy's; the encoder that takes them writes the oracle's payload
    and gap array bit for bit, for ragged and whole ranges, and its stream decodes."""
    cdf, perm = dist_spec(dist)
    for n in RANGE_SIZES:
        say("ranges", dist, "n =", n)
        d_dev = device_data(codec, n, 5 + n, cdf, perm)
        d = d_dev.cpu().numpy()
        h, rh = codec.histogram_ranges(d_dev)
        rows = int(codec._lib().hgap_range_count(n))
        assert rows == -(-n // RANGE)
        assert (codec.host_histogram(h).astype(np.int64) == np.bincount(d, minlength=256)).all()
        got = rh.cpu().numpy().view(np.uint32).reshape(-1, 256)[:rows].astype(np.int64)
        assert (got == range_counts(d, n)).all(), f"{dist} n={n}: range counts"
        hist = orc.histogram(d)
        ocb = orc.codebook(hist)
        pay, gap = orc.encode(d, ocb)
        # an explicit (default) shape runs the chained encoder at any size
        enc = codec.encode(d_dev, hist=hist, range_hist=rh, config={})
        assert (enc.payload_u32() == pay).all(), f"{dist} n={n}: payload"
        assert (enc.gap_u32() == gap).all(), f"{dist} n={n}: gap"
        # without one, small inputs take the per-tile look-back: the same stream
        enc2 = codec.encode(d_dev, hist=hist, range_hist=rh)
        assert (enc2.payload_u32() == pay).all() and (enc2.gap_u32() == gap).all()
        table = codec.DeviceTable(enc.codebook)
        out = codec.decode(enc.payload, enc.gap, enc.payload_words, enc.gap_elements, table, n)
        assert (out.cpu().numpy() == d).all(), f"{dist} n={n}: round trip"


def test_chained_encode_every_shape(codec, orc):
    """Every encoder shape (tiles of 4 KiB to 64 KiB, 64 down to 4 per range) writes the same
    stream through the chained encoder; repeated calls reuse one workspace."""
    cdf, perm = dist_spec("zipf")
    n = 3 * RANGE + 4097
    d_dev = device_data(codec, n, 21, cdf, perm)
    d = d_dev.cpu().numpy()
    hist = orc.histogram(d)
    pay, gap = orc.encode(d, orc.codebook(hist))
    _, rh = codec.histogram_ranges(d_dev)
    cb = codec.codebook_build(hist)
    sizes = codec.encode_sizes(cb, hist)
    ws = codec.Workspace()
    ran = 0
    for threads in (256, 512, 1024):
        for k in (16, 32, 64):
            cfg = {"enc_threads": threads, "enc_bytes_per_thread": k}
            say("chained shape", cfg)
            try:
                enc = codec.encode(d_dev, cb=cb, sizes=sizes, range_hist=rh, config=cfg,
                                   workspace=ws)
            except codec.HgapError as e:
                assert e.status == 8, cfg  # LDS image too large for this shape and codebook
                continue
            ran += 1
            assert (enc.payload_u32() == pay).all(), cfg
            assert (enc.gap_u32() == gap).all(), cfg
    assert ran >= 8


def check_encode(codec, orc, d_dev, label):
    d = d_dev.cpu().numpy()
    hist = orc.histogram(d)
    cb = codec.codebook_build(hist)
    enc = codec.encode(d_dev, cb=cb, hist=hist)
    ocb = orc.codebook(hist)
    pay, gap = orc.encode(d, ocb)
    assert enc.payload_words == pay.size, label
    gp = enc.payload_u32()
    gg = enc.gap_u32()
    if not (gp == pay).all():
        bad = np.nonzero(gp != pay)[0]
        pytest.fail(f"{label}: {bad.size} payload words differ, first at {bad[0]}")
    if not (gg == gap).all():
        bad = np.nonzero(gg != gap)[0]
        pytest.fail(f"{label}: {bad.size} gap words differ, first at {bad[0]}")
    return enc, ocb


SIZES = [1, 2, 31, 32, 33, 8191, 8192, 8193, 3 * 8192 + 17, (1 << 20) + 7]


@pytest.mark.parametrize("dist", ["zipf", "geometric", "skew", "uniform", "gauss", "k1", "k2", "fib"])
def test_encode_bit_exact_and_decode(codec, orc, dist):
    cdf, perm = dist_spec(dist)
    for n in SIZES:
        say("encode/decode", dist, "n =", n)
        d_dev = device_data(codec, n, 11 + n, cdf, perm)
        enc, ocb = check_encode(codec, orc, d_dev, f"{dist} n={n}")
        table = codec.DeviceTable(enc.codebook)
        out = codec.decode(enc.payload, enc.gap, enc.payload_words, enc.gap_elements, table, n)
        got = out.cpu().numpy()
        want = d_dev.cpu().numpy()
        if not (got == want).all():
            bad = np.nonzero(got != want)[0]
            pytest.fail(f"{dist} n={n}: {bad.size} bytes differ, first at {bad[0]}")


@pytest.mark.parametrize("prefix", [1, 4, 6, 9, 12, 13])
def test_decode_prefix_splits(codec, orc, prefix):
    """Every first-level width decodes the same bytes (one- and two-level tables); a table too
    large for LDS is refused with HGAP_ERR_UNSUPPORTED instead of being launched."""
    cdf, perm = dist_spec("fib")
    n = 300_001
    say("decode prefix", prefix)
    d_dev = device_data(codec, n, 5, cdf, perm)
    enc = codec.encode(d_dev)
    table = codec.DeviceTable(enc.codebook, prefix_bits=prefix)
    info = table.info
    assert (info.l2_entries == 0) == (prefix >= info.max_length)
    # the default entry point falls back to 256-segment tiles before refusing: refused when
    # the table leaves less than 16 bytes per segment (+ 48) of LDS for the output image
    image = codec._lib().hgap_decode_image_bytes(256, info.bytes // 4,
                                                 info.max_symbols_per_segment)
    need = (48 + 256 * info.max_symbols_per_segment + 15) // 16 * 16
    fixed = 4 * ((info.bytes // 4 + 3) // 4 * 4) + min(need, 16 * 256 + 48)
    if fixed <= 160 * 1024 - 256:  # static LDS of the kernels is under 256 bytes
        assert 0 < image <= need and image % 16 == 0
    elif fixed > 160 * 1024:
        assert image == 0
    if image == 0:
        with pytest.raises(codec.HgapError) as err:
            codec.decode(enc.payload, enc.gap, enc.payload_words, enc.gap_elements, table, n)
        assert err.value.status == 8
        return
    out = codec.decode(enc.payload, enc.gap, enc.payload_words, enc.gap_elements, table, n)
    assert bool((out == d_dev).all())


def test_decode_dense_tiles_through_windows(codec):
    """Half zero bytes (a 1-bit code: about 128 codewords per segment), half uniform bytes: the
    decoder sizes its LDS image for the stream's mean density, so the dense tiles go through the
    image in windows.  The output is the input."""
    import torch

    n = (3 << 20) + 11
    half = (n // 2) & ~15
    d = torch.zeros(n, dtype=torch.uint8, device="cuda")
    cdf, perm = dist_spec("uniform")
    codec.synth_fill(d[half:], 5, cdf, perm)
    enc = codec.encode(d)
    table = codec.DeviceTable(enc.codebook)
    info = table.info
    assert info.min_length == 1
    mean = n / enc.gap_elements
    expect = min(info.max_symbols_per_segment, -(-5 * n // (4 * enc.gap_elements)) + 2)
    image = codec._lib().hgap_decode_image_bytes(256, info.bytes // 4, expect)
    say("dense tiles: mean", round(mean, 1), "codewords/segment, image", image, "bytes")
    assert 0 < image < 256 * 100  # a tile of the zero half holds ~256 x 128 codewords
    out = codec.decode(enc.payload, enc.gap, enc.payload_words, enc.gap_elements, table, n)
    assert torch.equal(out[:n], d)


def test_repeated_calls_reuse_workspace(codec):
    """Look-back status words are cleared on every call: results repeat exactly."""
    import torch

    cdf, perm = dist_spec("zipf")
    n = (1 << 22) + 5
    d_dev = device_data(codec, n, 21, cdf, perm)
    ws = codec.Workspace()
    a = codec.encode(d_dev, workspace=ws)
    table = codec.DeviceTable(a.codebook)
    outs = []
    for _ in range(3):
        b = codec.encode(d_dev, cb=a.codebook, sizes=a.sizes, workspace=ws)
        assert torch.equal(a.payload[: a.payload_words], b.payload[: b.payload_words])
        assert torch.equal(a.gap[: a.sizes.gap_words], b.gap[: b.sizes.gap_words])
        outs.append(codec.decode(b.payload, b.gap, b.payload_words, b.gap_elements, table, n,
                                 workspace=ws))
    for o in outs:
        assert torch.equal(o, d_dev)


def test_stream_bytes_match_oracle_and_round_trip(codec, orc):
    for dist, n in (("zipf", 123_457), ("k1", 1000), ("geometric", 65_536)):
        say("stream", dist, n)
        cdf, perm = dist_spec(dist)
        d_dev = device_data(codec, n, 9, cdf, perm)
        d = d_dev.cpu().numpy()
        blob = codec.compress(d_dev)
        ocb = orc.codebook(orc.histogram(d))
        sz = orc.sizes(orc.histogram(d), ocb["length"])
        pay, gap = orc.encode(d, ocb, sz)
        assert blob == orc.stream(ocb, n, pay, gap, sz["payload_words"], sz["gap_elements"])
        assert (codec.decompress(blob) == d).all()
    assert codec.decompress(codec.compress(b"")).size == 0


@pytest.mark.slow
@pytest.mark.parametrize("name", list(synth.CONFIGS))
def test_full_size_configs(codec, orc, golden_codebooks, name):
    """Each benchmark configuration at its full size: histogram and codebook equal the golden
    (reference) ones, payload and gap array equal the oracle's bit for bit, decode returns the
    input byte for byte."""
    import torch

    n, seed, cdf, perm = synth.spec(name)
    say("full size", name, n)
    d_dev = device_data(codec, n, seed, cdf, perm)
    g = golden_codebooks
    i = list(g["tag"]).index(name)
    h = codec.host_histogram(codec.histogram(d_dev)).astype(np.uint64)
    assert (h == g["hist"][i]).all()
    enc = codec.encode(d_dev, hist=h)
    k = int(g["k"][i])
    assert enc.codebook.symbol_count == k
    assert list(enc.codebook.order_symbol[:k]) == list(g["file_symbol"][i][:k])
    assert list(enc.codebook.order_length[:k]) == list(g["file_length"][i][:k])
    table = codec.DeviceTable(enc.codebook)
    out = codec.decode(enc.payload, enc.gap, enc.payload_words, enc.gap_elements, table, n)
    assert torch.equal(out, d_dev), "round trip"
    say("  chained encode", name)
    h2, rh = codec.histogram_ranges(d_dev)
    assert (codec.host_histogram(h2).astype(np.uint64) == h).all()
    W, GW = enc.payload_words, int(enc.sizes.gap_words)
    enc2 = codec.encode(d_dev, cb=enc.codebook, sizes=enc.sizes, range_hist=rh, config={})
    assert torch.equal(enc2.payload[:W], enc.payload[:W]), "chained payload"
    assert torch.equal(enc2.gap[:GW], enc.gap[:GW]), "chained gap"
    del enc2
    # without a shape: the library's own choice (256 x 32, or 512 x 32 where its workgroups
    # divide the ranges more evenly; look-back below 64 MiB)
    enc3 = codec.encode(d_dev, cb=enc.codebook, sizes=enc.sizes, range_hist=rh)
    assert torch.equal(enc3.payload[:W], enc.payload[:W]), "default chained payload"
    assert torch.equal(enc3.gap[:GW], enc.gap[:GW]), "default chained gap"
    del enc3, rh
    say("  oracle encode", name)
    d = d_dev.cpu().numpy()
    del out
    ocb = orc.codebook(h)
    pay, gap = orc.encode(d, ocb)
    assert (enc.payload_u32() == pay).all(), "payload"
    assert (enc.gap_u32() == gap).all(), "gap"


# ---- reference-named entry points through the C ABI ---------------------------------------

class Hip:
    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]

    def malloc(self, n):
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), n) == 0
        return p

    def h2d(self, dst, arr):
        assert self.lib.hipMemcpy(dst, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0

    def d2h(self, arr, src):
        assert self.lib.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), src, arr.nbytes, 2) == 0


def test_reference_named_pipeline(cuda, capi, orc):
    """histgram -> store_symbols/qsort/boundary_PM -> encode -> get_table_info/get_twolevel_table
    -> decoder_l1_l2, called the way the reference's two main() functions call them."""
    from huffman_coding_gap_arrays_amd._capi import Codetable, Symbol, TableInfo

    hip = Hip()
    for dist, n in (("zipf", 200_003), ("fib", 50_001), ("k2", 999)):
        say("reference-named pipeline", dist, n)
        cdf, perm = dist_spec(dist)
        d = orc.synth(n, 0, 3, cdf, perm)
        words = (n + 3) // 4
        host_in = np.zeros(4 * (words + 1), np.uint8)
        host_in[:n] = d
        d_in = hip.malloc(4 * (words + 1))
        hip.h2d(d_in, host_in)
        d_hist = hip.malloc(1024)
        capi.histgram(d_hist, d_in, n)
        hist = np.zeros(256, np.uint32)
        hip.d2h(hist, d_hist)
        assert (hist == np.bincount(d, minlength=256)).all()
        syms = (Symbol * 256)()
        k = capi.store_symbols(hist.ctypes.data_as(ctypes.c_void_p), syms)
        order = sorted(range(k), key=lambda j: syms[j].num)
        srt = (Symbol * 256)()
        for j, o in enumerate(order):
            srt[j] = syms[o]
        table = (Codetable * 256)()
        capi.boundary_PM(srt, k, table)
        bits = capi.get_outputfilesize(srt, k)
        G = (bits + 127) // 128
        W = (bits + 31) // 32
        out = np.zeros(W + (G + 7) // 8, np.uint32)
        capi.encode(out.ctypes.data_as(ctypes.c_void_p), W, d_in, n, G, table)  # frees d_in
        ocb = orc.codebook(hist.astype(np.uint64))
        pay, gap = orc.encode(d, ocb)
        assert (out[:W] == pay).all() and (out[W:] == gap).all()
        fsyms = (Symbol * 256)()
        for j in range(k):
            fsyms[j] = srt[k - 1 - j]
        info = TableInfo()
        P = capi.get_table_info(fsyms, k, 10, ctypes.byref(info))
        size = 4 * info.ptrtable_size + 256 + info.l1table_size + info.l2table_size
        stage = np.zeros(size, np.uint8)
        capi.get_twolevel_table(stage.ctypes.data_as(ctypes.c_void_p), P, fsyms, k, info)
        inp = np.concatenate([gap, pay]).astype(np.uint32)
        res = np.zeros(4 * words, np.uint8)
        capi.decoder_l1_l2(inp.ctypes.data_as(ctypes.c_void_p), W,
                           res.ctypes.data_as(ctypes.c_void_p), n, G,
                           stage.ctypes.data_as(ctypes.c_void_p), size, P, k, info)
        assert (res[:n] == d).all()
        hip.lib.hipFree(d_hist)


def test_cli_tools_round_trip(cuda, orc, tmp_path):
    """hgap_encoder / hgap_decoder: the written file is the oracle's stream byte for byte and
    decodes back to the input (reference usage: encoder in out; decoder in [out])."""
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    bindir = os.path.join(root, "huffman_coding_gap_arrays_amd", "bin")
    cdf, perm = dist_spec("geometric")
    d = orc.synth(2_000_003, 0, 17, cdf, perm)
    src = tmp_path / "in.bin"
    src.write_bytes(d.tobytes())
    say("cli encode")
    r = subprocess.run([os.path.join(bindir, "hgap_encoder"), str(src), str(tmp_path / "c.huff")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    blob = (tmp_path / "c.huff").read_bytes()
    h = orc.histogram(d)
    ocb = orc.codebook(h)
    sz = orc.sizes(h, ocb["length"])
    pay, gap = orc.encode(d, ocb, sz)
    assert blob == orc.stream(ocb, d.size, pay, gap, sz["payload_words"], sz["gap_elements"])
    say("cli decode")
    r = subprocess.run([os.path.join(bindir, "hgap_decoder"), str(tmp_path / "c.huff")],
                       capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "decodedfile").read_bytes() == d.tobytes()
    r = subprocess.run([os.path.join(bindir, "hgap_decoder"), str(tmp_path / "missing")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr


@pytest.mark.slow
def test_largest_stream_size(codec):
    """n = 2^32 - 1, the largest size the stream header's u32 field holds: histogram counts sum
    to n, sizes are consistent, and the decode returns the input byte for byte (a size-free
    property; the oracle would take minutes at this size)."""
    import torch

    n = (1 << 32) - 1
    say("largest stream size", n)
    cdf, perm = dist_spec("skew")
    d_dev = device_data(codec, n, 17, cdf, perm)
    h = codec.host_histogram(codec.histogram(d_dev)).astype(np.uint64)
    assert int(h.sum()) == n and int(h.max()) >= 1 << 31  # byte 0: ~0.9 n, past int32
    enc = codec.encode(d_dev, hist=h)
    sz = enc.sizes
    assert sz.payload_words == (int(sz.bits) + 31) // 32 and sz.gap_elements == (int(sz.bits) + 127) // 128
    table = codec.DeviceTable(enc.codebook)
    out = codec.decode(enc.payload, enc.gap, enc.payload_words, enc.gap_elements, table, n)
    assert torch.equal(out[:n], d_dev)
    del out, enc, d_dev
    torch.cuda.empty_cache()

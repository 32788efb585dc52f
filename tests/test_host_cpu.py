"""Host-side parts of libhgap.so (codebook, sizes, decode tables, stream header, reference-named
entry points) against the oracle and the reference's golden vectors.  CPU only: none of these
calls touches a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from huffman_coding_gap_arrays_amd import _capi, synth
from huffman_coding_gap_arrays_amd._capi import (Codebook, Codetable, DectableInfo, Sizes,
                                                 Symbol, TableInfo)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def build_cb(capi, hist):
    h = np.ascontiguousarray(hist, dtype=np.uint32)
    cb = Codebook()
    assert capi.hgap_codebook_build(vp(h), ctypes.byref(cb)) == 0
    return cb


def cb_arrays(cb):
    k = cb.symbol_count
    return (k, np.frombuffer(bytes(cb.order_symbol), np.uint8)[:k],
            np.frombuffer(bytes(cb.order_length), np.uint8)[:k],
            np.array(cb.code[:], np.uint32), np.array(cb.length[:], np.uint32))


def test_library_exports_every_declared_symbol(capi):
    names = set()
    for hdr in ("hgap.h", "hgap_compat.h"):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = text.split("#ifdef __cplusplus\n}")[0]  # C++-only inline overloads follow
        for m in re.finditer(r"^[A-Za-z_][\w \t\*]*?\b([a-z_][a-zA-Z0-9_]*)\s*\(", text, re.M):
            if m.group(1) not in ("if", "return", "sizeof", "defined"):
                names.add(m.group(1))
    assert {"hgap_encode", "hgap_decode", "histgram", "encode", "decoder_l1_l2",
            "boundary_PM", "get_twolevel_table"} <= names
    for name in sorted(names):
        assert hasattr(capi, name), name
        assert name in _capi.PROTOTYPES, name


def test_codebook_build_matches_reference_golden(capi, golden_codebooks):
    g = golden_codebooks
    for i in range(len(g["k"])):
        if int(g["hist"][i].max()) >= 1 << 32:
            continue
        k, osym, olen, code, length = cb_arrays(build_cb(capi, g["hist"][i]))
        kk = int(g["k"][i])
        assert k == kk, g["tag"][i]
        assert (osym == g["file_symbol"][i][:kk]).all() and (olen == g["file_length"][i][:kk]).all()
        assert (code == g["code"][i]).all() and (length == g["length"][i]).all(), g["tag"][i]


def test_codebook_empty_and_single(capi):
    k, *_ = cb_arrays(build_cb(capi, np.zeros(256, np.uint32)))
    assert k == 0
    h = np.zeros(256, np.uint32)
    h[200] = 5
    k, osym, olen, code, length = cb_arrays(build_cb(capi, h))
    assert (k, osym[0], olen[0], code[200], length[200]) == (1, 200, 1, 0, 1)


def test_encode_sizes_match_oracle(capi, orc, golden_codebooks):
    g = golden_codebooks
    for i in range(0, len(g["k"]), 7):
        h = g["hist"][i]
        if int(h.max()) >= 1 << 32 or int(h.sum()) >= 1 << 32:
            continue
        cb = build_cb(capi, h)
        sz = Sizes()
        assert capi.hgap_encode_sizes(ctypes.byref(cb), vp(np.ascontiguousarray(h, np.uint32)),
                                      ctypes.byref(sz)) == 0
        o = orc.sizes(h, np.array(cb.length[:], np.uint32))
        assert (sz.bits, sz.payload_words, sz.gap_elements, sz.gap_words) == \
            (o["bits"], o["payload_words"], o["gap_elements"], o["gap_words"])


def test_codebook_from_order_rejects_invalid(capi):
    cb = Codebook()

    def build(sym, lens):
        s = np.array(sym, np.uint8)
        l = np.array(lens, np.uint8)
        return capi.hgap_codebook_from_order(vp(s), vp(l), len(sym), ctypes.byref(cb))

    assert build([1, 2, 3], [1, 2, 2]) == 0
    assert build([1, 2, 3], [1, 2, 17]) == 3        # length > 16
    assert build([1, 2, 3], [2, 1, 2]) == 3         # not monotone
    assert build([1, 1, 3], [1, 2, 2]) == 3         # duplicate symbol
    assert build([1, 2, 3, 4], [1, 1, 2, 2]) == 3   # over-subscribed


def walk_native_table(image, info, payload, gap, W, G, n, check_second=True):
    """Decodes with a libhgap table image in Python, by the format's rules: segment starts from
    the gap array; every step starts a codeword before the segment limit; a direct entry is
    sym1 | bits << 8 | sym2 << 16 | symbols << 24; a two-codeword entry whose second codeword
    (length from the image's trailing length block) starts at or past the limit keeps only its
    first; pointer entries (bit 31) lead to the second level.  check_second=False plants the
    mistake of keeping the second codeword always."""
    P = info.prefix_bits
    lengths = image[info.l1_entries + info.l2_entries:].view(np.uint8)
    assert image.size == info.l1_entries + info.l2_entries + 64
    bits = np.unpackbits(payload.astype(">u4").view(np.uint8))
    bits = np.concatenate([bits, np.zeros(64, np.uint8)])

    def lookup(p):
        win = int("".join(map(str, bits[p:p + 32])), 2)
        e = int(image[win >> (32 - P)])
        if e >> 31:
            width = (e >> 24) & 31
            e = int(image[info.l1_entries + (e & 0xFFFFFF) + (((win << P) & 0xFFFFFFFF) >> (32 - width))])
        return e

    out = []
    for s in range(G):
        start = 0 if s == 0 else (int(gap[(s - 1) // 8]) >> (4 * ((s - 1) % 8))) & 15
        p = 128 * s + start
        end = min(128 * (s + 1), 32 * W)
        while p < end:
            e = lookup(p)
            assert e >> 31 == 0 and (e >> 24) in (1, 2)
            used = (e >> 8) & 0xFF
            out.append(e & 0xFF)
            if e >> 24 == 2:
                len2 = int(lengths[(e >> 16) & 0xFF])
                assert 1 <= len2 < used
                if p + used - len2 < end or not check_second:
                    out.append((e >> 16) & 0xFF)
            p += used
    return np.array(out[:n], np.uint8)


@pytest.mark.parametrize("prefix", [0, 4, 6, 8, 11, 16])
@pytest.mark.parametrize("dist", ["zipf", "geometric", "fib", "k1", "uniform"])
def test_dectable_decodes_oracle_streams(capi, orc, prefix, dist):
    rng = np.random.default_rng(3)
    n = 2500
    if dist == "zipf":
        d = synth.fill_numpy(n, 0, 5, synth.cdf_from_pmf(synth.pmf_zipf(1.1)), synth.permutation(3))
    elif dist == "geometric":
        d = synth.fill_numpy(n, 0, 6, synth.cdf_from_pmf(synth.pmf_geometric(0.1)), synth.IDENTITY)
    elif dist == "fib":
        fib = [1, 1]
        while len(fib) < 26:
            fib.append(fib[-1] + fib[-2])
        w = np.array(fib[:24], float)
        d = np.concatenate([np.arange(24, dtype=np.uint8),
                            rng.choice(np.arange(24, dtype=np.uint8), n, p=w / w.sum())])
    elif dist == "k1":
        d = np.full(n, 9, np.uint8)
    else:
        d = rng.integers(0, 256, n, dtype=np.uint8)
    h = orc.histogram(d)
    cb = build_cb(capi, h.astype(np.uint32))
    info = DectableInfo()
    assert capi.hgap_dectable_build(ctypes.byref(cb), prefix, ctypes.byref(info), None, 0) == 0
    img = np.zeros(info.bytes // 4, np.uint32)
    assert capi.hgap_dectable_build(ctypes.byref(cb), prefix, ctypes.byref(info), vp(img),
                                    img.nbytes) == 0
    ocb = orc.codebook(h)
    sz = orc.sizes(h, ocb["length"])
    pay, gap = orc.encode(d, ocb, sz)
    out = walk_native_table(img, info, pay, gap, sz["payload_words"], sz["gap_elements"], d.size)
    assert (out == d).all()
    # trailing block: the oracle's code length of every byte value (0 when absent)
    tail = img[info.l1_entries + info.l2_entries:].view(np.uint8)
    assert (tail == np.asarray(ocb["length"], np.uint8)).all()
    assert info.max_symbols_per_segment == -(-128 // info.min_length)
    if dist == "zipf" and prefix in (0, 11):
        # keeping a second codeword that starts past the segment limit duplicates symbols
        bad = walk_native_table(img, info, pay, gap, sz["payload_words"], sz["gap_elements"],
                                d.size, check_second=False)
        assert not (bad == d).all()


def test_stream_header_round_trip_and_layout(capi, orc):
    d = synth.fill_numpy(9000, 0, 4, synth.cdf_from_pmf(synth.pmf_zipf(1.1)), synth.permutation(5))
    h = orc.histogram(d).astype(np.uint32)
    cb = build_cb(capi, h)
    sz = Sizes()
    assert capi.hgap_encode_sizes(ctypes.byref(cb), vp(h), ctypes.byref(sz)) == 0
    hdr = np.zeros(int(capi.hgap_stream_header_bytes(cb.symbol_count)), np.uint8)
    wrote = ctypes.c_size_t()
    assert capi.hgap_stream_write_header(ctypes.byref(cb), d.size, ctypes.byref(sz), vp(hdr),
                                         hdr.size, ctypes.byref(wrote)) == 0
    ocb = orc.codebook(h)
    pay, gap = orc.encode(d, ocb)
    blob = orc.stream(ocb, d.size, pay, gap, sz.payload_words, sz.gap_elements)
    assert bytes(hdr) == blob[: hdr.size]
    assert capi.hgap_stream_bytes(cb.symbol_count, sz.payload_words, sz.gap_elements) == len(blob)
    buf = np.frombuffer(blob, np.uint8).copy()
    cb2, n2, sz2 = Codebook(), ctypes.c_uint32(), Sizes()
    go, po = ctypes.c_uint64(), ctypes.c_uint64()
    assert capi.hgap_stream_parse_header(vp(buf), buf.size, ctypes.byref(cb2), ctypes.byref(n2),
                                         ctypes.byref(sz2), ctypes.byref(go), ctypes.byref(po)) == 0
    assert n2.value == d.size and sz2.payload_words == sz.payload_words
    assert sz2.gap_elements == sz.gap_elements and go.value == hdr.size
    assert po.value == hdr.size + 4 * sz.gap_words
    assert list(cb2.code) == list(cb.code) and list(cb2.length) == list(cb.length)
    # truncated or inconsistent streams are rejected with HGAP_ERR_FORMAT
    for cut in (0, 7, hdr.size - 1, len(blob) - 1):
        assert capi.hgap_stream_parse_header(vp(buf), cut, ctypes.byref(cb2), ctypes.byref(n2),
                                             ctypes.byref(sz2), None, None) == 6
    bad = buf.copy()
    bad[8 + 2 * cb.symbol_count + 8] ^= 1  # G no longer matches W
    assert capi.hgap_stream_parse_header(vp(bad), bad.size, ctypes.byref(cb2), ctypes.byref(n2),
                                         ctypes.byref(sz2), None, None) == 6


def test_device_entry_points_validate_before_launch(capi):
    """Argument errors are returned before any device work (no GPU needed)."""
    cb, sz = Codebook(), Sizes()
    sz.payload_words, sz.gap_elements, sz.gap_words, sz.bits = 4, 1, 1, 128
    assert capi.hgap_encode(None, 100, ctypes.byref(cb), ctypes.byref(sz), None, None, None, 0,
                            None) == 1
    assert capi.hgap_encode(ctypes.c_void_p(16), 100, ctypes.byref(cb), ctypes.byref(sz),
                            ctypes.c_void_p(32), ctypes.c_void_p(48), ctypes.c_void_p(64), 8,
                            None) == 4  # workspace too small
    assert capi.hgap_encode(ctypes.c_void_p(17), 100, ctypes.byref(cb), ctypes.byref(sz),
                            ctypes.c_void_p(32), ctypes.c_void_p(48), ctypes.c_void_p(64), 4096,
                            None) == 1  # misaligned input
    info = DectableInfo()
    p = ctypes.c_void_p(16)
    assert capi.hgap_decode(p, 8, p, 2, p, ctypes.byref(info), p, 10, p, 4096, None) == 3
    info.symbol_count, info.prefix_bits, info.l1_entries, info.min_length = 2, 1, 2, 1
    info.bytes = 4 * 2  # no trailing length block: not an image of this library
    assert capi.hgap_decode(p, 8, p, 2, p, ctypes.byref(info), p, 10, p, 4096, None) == 3
    info.bytes = 4 * (2 + 64)
    assert capi.hgap_decode(p, 8, p, 3, p, ctypes.byref(info), p, 10, p, 4096, None) == 2
    assert capi.hgap_decode(p, 8, p, 2, p, ctypes.byref(info), p, 10, p, 8, None) == 4
    assert capi.hgap_histogram(ctypes.c_void_p(3), 10, ctypes.c_void_p(16), None) == 1
    # per-range histograms and the encoder that takes them
    R = 256 * 1024
    assert [capi.hgap_range_count(x) for x in (0, 1, R, R + 1, (1 << 32) - 1)] == \
        [0, 1, 1, 2, 1 << 14]
    q = ctypes.c_void_p(16)
    assert capi.hgap_histogram_ranges(q, 10, None, q, None) == 1
    assert capi.hgap_histogram_ranges(q, 10, q, None, None) == 1
    assert capi.hgap_histogram_ranges(q, 10, q, ctypes.c_void_p(20), None) == 1
    assert capi.hgap_histogram_ranges(q, 1 << 32, q, q, None) == 2
    assert capi.hgap_encode_ranges(q, 100, ctypes.byref(cb), ctypes.byref(sz), None, q, q, q,
                                   4096, None, None) == 1  # no range counts
    assert capi.hgap_encode_ranges(q, 100, ctypes.byref(cb), ctypes.byref(sz),
                                   ctypes.c_void_p(24), q, q, q, 4096, None, None) == 1
    assert capi.hgap_encode_ranges(q, 100, ctypes.byref(cb), ctypes.byref(sz), q, q, q, q, 8,
                                   None, None) == 4  # workspace too small
    # batched per-range counts and the chained batched encoder
    assert capi.hgap_batch_histogram_ranges(q, 1, 5, None, None) == 1
    assert capi.hgap_batch_histogram_ranges(q, 1, 5, ctypes.c_void_p(20), None) == 1
    assert capi.hgap_batch_encode_ranges_workspace_bytes(10) == 16 + 10 * 12
    assert capi.hgap_batch_encode_ranges(q, 1, 10, 8, q, None, q, 4096, None, None) == 1
    assert capi.hgap_batch_encode_ranges(q, 1, 10, 0, q, q, q, 4096, None, None) == 3
    assert capi.hgap_batch_encode_ranges(q, 1, 10, 8, q, q, q, 135, None, None) == 4
    assert capi.hgap_encode_workspace_bytes(0) == 16
    assert capi.hgap_encode_workspace_bytes(8193) == 16 + 3 * 8  # smallest tile: 4 KiB
    # ticket + status word of one offsets block, one count byte per segment (whole 1024-segment
    # tiles), one sum per 64 of those, one offset per 256-segment tile (the smallest tile)
    assert capi.hgap_decode_workspace_bytes(257) == 16 + 16 + 1024 + 64 + 2 * 8
    # launch shapes outside the supported set are refused before anything is launched
    from huffman_coding_gap_arrays_amd._capi import LaunchConfig, default_config

    good = default_config()
    assert good.enc_threads in (256, 512, 1024) and good.lookback_window <= 256
    for field, value in (("enc_threads", 300), ("enc_bytes_per_thread", 8), ("dec_threads", 128),
                         ("lookback_window", 0), ("lookback_window", 512), ("prefix_bits", 17)):
        bad = LaunchConfig.from_dict({**good.as_dict(), field: value})
        assert capi.hgap_encode_cfg(ctypes.c_void_p(16), 100, ctypes.byref(cb), ctypes.byref(sz),
                                    ctypes.c_void_p(32), ctypes.c_void_p(48),
                                    ctypes.c_void_p(64), 1 << 20, ctypes.byref(bad), None) == 1
        assert capi.hgap_decode_cfg(p, 8, p, 2, p, ctypes.byref(info), p, 10, p, 4096,
                                    ctypes.byref(bad), None) == 1


# ---- reference-named entry points (include/hgap_compat.h) ---------------------------------

def ref_symbols(capi, hist):
    num = np.ascontiguousarray(hist, np.uint32)
    syms = (Symbol * 256)()
    k = capi.store_symbols(vp(num), syms)
    return k, syms


def test_compat_store_symbols_boundary_pm_outputsize(capi, golden_codebooks):
    g = golden_codebooks
    for i in range(len(g["k"])):
        h = g["hist"][i]
        if int(h.max()) >= 1 << 32:
            continue
        k, syms = ref_symbols(capi, h)
        assert k == int(g["k"][i])
        present = np.nonzero(h)[0]
        assert [syms[j].symbol for j in range(k)] == list(present)  # byte order
        order = sorted(range(k), key=lambda j: syms[j].num)  # stable, like glibc qsort
        sorted_syms = (Symbol * 256)()
        for j, o in enumerate(order):
            sorted_syms[j] = syms[o]
        table = (Codetable * 256)()
        capi.boundary_PM(sorted_syms, k, table)
        assert [table[b].code for b in range(256)] == list(g["code"][i])
        assert [table[b].length for b in range(256)] == list(g["length"][i])
        bits = capi.get_outputfilesize(sorted_syms, k)
        assert bits == int((h.astype(np.uint64) * g["length"][i].astype(np.uint64)).sum())


def walk_stage_table(stage, P, info, payload, gap, W, G, n):
    """Decodes with a stage table in the reference layout (decoder/src/decoder.cu:531-555
    semantics, boundary = l1table_size), in Python."""
    ptr = stage[: 4 * info.ptrtable_size].view("<u4")
    lens = stage[4 * info.ptrtable_size: 4 * info.ptrtable_size + 256]
    l1 = stage[4 * info.ptrtable_size + 256:]
    l2 = l1[info.l1table_size:]
    bits = np.unpackbits(payload.astype(">u4").view(np.uint8))
    bits = np.concatenate([bits, np.zeros(64, np.uint8)])
    out = []
    for s in range(G):
        start = 0 if s == 0 else (int(gap[(s - 1) // 8]) >> (4 * ((s - 1) % 8))) & 15
        p = 128 * s + start
        end = min(128 * (s + 1), 32 * W)
        while p < end:
            win = int("".join(map(str, bits[p:p + 32])), 2)
            dc = win >> (32 - P)
            if dc < info.l1table_size:
                sym = int(l1[dc])
            else:
                v = int(ptr[dc - info.l1table_size])
                width, off = v >> 16, v & 0xFFFF
                sym = int(l2[off + ((((win << P) & 0xFFFFFFFF)) >> (32 - width))])
            out.append(sym)
            p += int(lens[sym])
    return np.array(out[:n], np.uint8)


def test_compat_two_level_table(capi, orc, golden_codebooks, golden_tables):
    g, t = golden_codebooks, golden_tables
    rng = np.random.default_rng(0)
    checked = 0
    for i in range(len(g["k"])):
        k = int(g["k"][i])
        if k < 2 or int(g["hist"][i].max()) >= 1 << 32:
            continue
        syms = (Symbol * 256)()
        for j in range(k):
            syms[j].symbol = int(g["file_symbol"][i][j])
            syms[j].length = int(g["file_length"][i][j])
        info = TableInfo()
        P = capi.get_table_info(syms, k, 10, ctypes.byref(info))
        assert P == int(t["prefix_bit"][i])
        # exact sizes from the code structure: first long prefix, one block per distinct
        # 10-bit prefix of long codes, block width = its longest code - 10
        lens = g["file_length"][i][:k].astype(int)
        codes = [int(g["code"][i][s]) for s in g["file_symbol"][i][:k]]
        groups = {}
        for c, l in zip(codes, lens):
            if l > 10:
                groups[c >> (l - 10)] = max(groups.get(c >> (l - 10), 0), l - 10)
        if groups:
            assert info.l1table_size == min(groups) == int(t["boundary"][i])
            assert info.ptrtable_size == len(groups)
            assert info.l2table_size == sum(1 << w for w in groups.values())
            # the reference's get_table_info reads the next symbol's code (get_table.cpp:20-27):
            # its first-level size is never below the true boundary
            assert int(t["sizes"][i][0]) >= info.l1table_size
        else:
            assert (info.l1table_size, info.l2table_size, info.ptrtable_size) == \
                tuple(int(x) for x in t["sizes"][i])
        size = 4 * info.ptrtable_size + 256 + info.l1table_size + info.l2table_size
        stage = np.zeros(size, np.uint8)
        boundary = capi.get_twolevel_table(vp(stage), P, syms, k, info)
        if info.ptrtable_size:
            assert boundary == info.l1table_size
            if t["unwritten_l1"][i] == 0:
                assert boundary == int(t["boundary"][i])
        if i % 9 == 0 and checked < 20:
            checked += 1
            present = np.nonzero(g["hist"][i])[0]
            d = rng.choice(present.astype(np.uint8), 400)
            d = np.concatenate([present.astype(np.uint8), d])
            cb = dict(k=k, code=g["code"][i], length=g["length"][i])
            bits = int(g["length"][i][d].sum())
            W, G = (bits + 31) // 32, (bits + 127) // 128
            pay, gap = orc.encode(d, cb, dict(payload_words=W, gap_elements=G))
            assert (walk_stage_table(stage, P, info, pay, gap, W, G, d.size) == d).all()
    assert checked >= 10


def test_python_wrappers_import_and_build_tables_on_cpu():
    """The torch-backed wrapper module imports and its host-side helpers work without a GPU."""
    import huffman_coding_gap_arrays_amd as pkg
    from huffman_coding_gap_arrays_amd import codec

    assert pkg.codec is codec
    cb = codec.codebook_build(np.arange(256) + 1)
    assert cb.symbol_count == 256
    t = codec.DeviceTable(cb, device="cpu")
    assert t.info.l1_entries == 1 << t.info.prefix_bits and t.table.numel() * 4 >= t.info.bytes
    blob = codec.stream_bytes(codec.Encoded(0, codec.Codebook(), codec.Sizes(),
                                            *[__import__("torch").zeros(4, dtype=__import__("torch").int32)] * 2))
    cb2, n, sz, gap, pay = codec.parse_stream(blob)
    assert n == 0 and cb2.symbol_count == 0 and gap.size == 0 and pay.size == 0


def test_batch_job_layout_and_plans(capi, golden_codebooks):
    """Job descriptors have the C layout; the plan functions lay out tiles back to back."""
    from huffman_coding_gap_arrays_amd import batch
    from huffman_coding_gap_arrays_amd._capi import Codebook, LaunchConfig, default_config

    assert (batch.HIST_JOB.itemsize, batch.ENC_JOB.itemsize, batch.DEC_JOB.itemsize) == (40, 64, 80)
    rng = np.random.default_rng(2)
    ns = rng.integers(0, 3 << 20, 50)
    ns[[3, 17]] = 0  # empty buffers own no ranges
    jobs = np.zeros(50, batch.HIST_JOB)
    jobs["n"] = ns
    total = capi.hgap_batch_plan_histogram(vp(jobs), 50)
    assert total == ns.sum() and (jobs["base"] == np.concatenate([[0], np.cumsum(ns)[:-1]])).all()
    rows = -(-ns // (256 * 1024))
    assert (jobs["range_base"] == np.concatenate([[0], np.cumsum(rows)[:-1]])).all()
    ej = np.zeros(50, batch.ENC_JOB)
    ej["n"] = ns
    ranges = capi.hgap_batch_plan_encode_ranges(vp(ej), 50, None)  # default 32 KiB tiles
    assert ranges == rows.sum() and (ej["base"] == jobs["range_base"]).all()
    assert (ej["num_tiles"] == -(-ns // 32768)).all()
    cfg = LaunchConfig.from_dict({**default_config().as_dict(), "enc_threads": 256,
                                  "enc_bytes_per_thread": 32, "dec_threads": 512})
    ej = np.zeros(50, batch.ENC_JOB)
    ej["n"] = ns
    tiles = capi.hgap_batch_plan_encode(vp(ej), 50, ctypes.byref(cfg))
    want = -(-ns // 8192)
    assert tiles == want.sum() and (ej["num_tiles"] == want).all()
    assert (ej["base"] == np.concatenate([[0], np.cumsum(want)[:-1]])).all()
    dj = np.zeros(50, batch.DEC_JOB)
    dj["n"] = ns
    dj["gap_elements"] = ns // 3
    tiles = capi.hgap_batch_plan_decode(vp(dj), 50, ctypes.byref(cfg))
    want = np.where(ns > 0, -(-(ns // 3) // 512), 0)
    assert tiles == want.sum() and (dj["num_tiles"] == want).all()
    # threaded codebook builds equal the one-at-a-time ones
    g = golden_codebooks
    keep = [i for i in range(len(g["k"])) if int(g["hist"][i].max()) < 1 << 32]
    h = np.ascontiguousarray(g["hist"][keep], dtype=np.uint32)
    cbs = batch.codebooks(h, threads=8)
    for j, i in enumerate(keep):
        assert list(cbs[j].code) == list(g["code"][i]) and list(cbs[j].length) == list(g["length"][i])


def test_batch_prepare_and_table_batch_match_single_calls(capi, golden_codebooks):
    from huffman_coding_gap_arrays_amd import _capi, batch

    def vp(a):
        return a.ctypes.data_as(ctypes.c_void_p)

    g = golden_codebooks
    # histograms whose streams fit the format's 32-bit sizes (the others report HGAP_ERR_SIZE)
    keep = [i for i in range(len(g["k"]))
            if int(g["hist"][i].max()) < 1 << 32 and int(g["bits"][i]) < 1 << 36]
    h = np.ascontiguousarray(g["hist"][keep], dtype=np.uint32)
    B = len(keep)
    assert B > 300
    cbs = (_capi.Codebook * B)()
    sizes = (_capi.Sizes * B)()
    tables = np.zeros((B, 256), np.uint32)
    assert capi.hgap_batch_prepare(vp(h), B, cbs, sizes, vp(tables), 8) == 0
    for j in range(B):
        cb = _capi.Codebook()
        assert capi.hgap_codebook_build(vp(h[j]), ctypes.byref(cb)) == 0
        assert list(cb.code) == list(cbs[j].code) and list(cb.length) == list(cbs[j].length)
        sz = _capi.Sizes()
        assert capi.hgap_encode_sizes(ctypes.byref(cb), vp(h[j]), ctypes.byref(sz)) == 0
        assert (sz.bits, sz.payload_words, sz.gap_words) == \
            (sizes[j].bits, sizes[j].payload_words, sizes[j].gap_words)
        t = np.zeros(256, np.uint32)
        assert capi.hgap_encode_table(ctypes.byref(cb), vp(t)) == 0
        assert (t == tables[j]).all()
    # decode tables: infos, then images at caller-chosen offsets, equal to one-at-a-time builds
    prefix = np.array([(0, 4, 8, 11, 12)[j % 5] for j in range(B)], np.uint32)
    infos = (_capi.DectableInfo * B)()
    assert capi.hgap_dectable_build_batch(cbs, B, vp(prefix), infos, None, None, 0, 8) == 0
    nbytes = np.array([infos[j].bytes for j in range(B)], np.int64)
    offs = np.zeros(B, np.uint64)
    offs[1:] = np.cumsum((nbytes + 255) // 256 * 256)[:-1]
    arena = np.zeros(int(offs[-1]) + int(nbytes[-1]) + 64, np.uint8)
    assert capi.hgap_dectable_build_batch(cbs, B, vp(prefix), infos, vp(arena), vp(offs),
                                          arena.nbytes, 8) == 0
    for j in range(B):
        info = _capi.DectableInfo()
        img = np.zeros(max(1, int(nbytes[j]) // 4), np.uint32)
        assert capi.hgap_dectable_build(ctypes.byref(cbs[j]), int(prefix[j]), ctypes.byref(info),
                                        vp(img), img.nbytes) == 0
        got = arena[int(offs[j]):int(offs[j]) + int(nbytes[j])].view(np.uint32)
        assert (got == img[:len(got)]).all() and info.l1_entries == infos[j].l1_entries
    # an arena too small is reported, not overrun
    assert capi.hgap_dectable_build_batch(cbs, B, vp(prefix), infos, vp(arena), vp(offs),
                                          int(offs[-1]), 8) == 7
    assert batch.INFO.itemsize == ctypes.sizeof(_capi.DectableInfo)

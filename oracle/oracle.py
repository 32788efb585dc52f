"""ctypes binding of the oracle (oracle/liboracle.so, built from oracle/hgap_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker the HIP product is compared against.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_ENC = os.path.join(_HERE, "_ref", "libref_encoder.so")
REF_DEC = os.path.join(_HERE, "_ref", "libref_decoder.so")

_u8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32 = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64 = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "oracle"], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = ctypes.CDLL(LIB)
    L.orc_histogram.argtypes = [_u8, ctypes.c_uint64, _u64]
    L.orc_sort_symbols.argtypes = [_u64, _u8, _u64]
    L.orc_sort_symbols.restype = ctypes.c_int
    L.orc_boundary_pm.argtypes = [_u64, ctypes.c_int, _u8, ctypes.POINTER(ctypes.c_int)]
    L.orc_boundary_pm.restype = ctypes.c_int
    L.orc_build_codebook.argtypes = [_u64, _u32, _u32, _u8, _u8, ctypes.POINTER(ctypes.c_int)]
    L.orc_build_codebook.restype = ctypes.c_int
    L.orc_sizes.argtypes = [_u64, _u32] + [ctypes.POINTER(ctypes.c_uint64)] * 4
    L.orc_encode.argtypes = [_u8, ctypes.c_uint64, _u32, _u32, _u32, ctypes.c_uint64, _u32,
                             ctypes.c_uint64]
    L.orc_encode.restype = ctypes.c_int
    L.orc_decode_segments.argtypes = [_u32, ctypes.c_uint64, _u32, ctypes.c_uint64, _u8, _u8,
                                      ctypes.c_int, _u8, ctypes.c_uint64, ctypes.c_void_p]
    L.orc_decode_segments.restype = ctypes.c_int64
    L.orc_stream_size.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
    L.orc_stream_size.restype = ctypes.c_int64
    L.orc_stream_write.argtypes = [_u8, ctypes.c_uint64, ctypes.c_int, _u8, _u8, ctypes.c_uint32,
                                   ctypes.c_uint32, ctypes.c_uint32, _u32, _u32]
    L.orc_stream_write.restype = ctypes.c_int64
    L.orc_histogram_mt.argtypes = [_u8, ctypes.c_uint64, _u64, ctypes.c_int]
    L.orc_encode_mt.argtypes = [_u8, ctypes.c_uint64, _u32, _u32, _u32, ctypes.c_uint64, _u32,
                                ctypes.c_uint64, ctypes.c_int]
    L.orc_encode_mt.restype = ctypes.c_int
    L.orc_decode_segments_mt.argtypes = [_u32, ctypes.c_uint64, _u32, ctypes.c_uint64, _u8, _u8,
                                         ctypes.c_int, _u8, ctypes.c_uint64, ctypes.c_int]
    L.orc_decode_segments_mt.restype = ctypes.c_int64
    L.orc_splitmix64_at.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.orc_splitmix64_at.restype = ctypes.c_uint64
    L.orc_synth_fill.argtypes = [_u8, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, _u64,
                                 _u8]
    _lib = L
    return L


def histogram(data):
    d = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    h = np.zeros(256, np.uint64)
    load().orc_histogram(d, d.size, h)
    return h


def histogram_mt(data, threads=0):
    d = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    h = np.zeros(256, np.uint64)
    load().orc_histogram_mt(d, d.size, h, threads)
    return h


def encode_mt(data, cb, sz, threads=0):
    """orc_encode_mt: same outputs as encode(), on `threads` cores (0 = all)."""
    d = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    W, G = sz["payload_words"], sz["gap_elements"]
    pay = np.empty(max(W, 1), np.uint32)
    gap = np.empty(max((G + 7) // 8, 1), np.uint32)
    rc = load().orc_encode_mt(d, d.size, cb["code"], cb["length"], pay, W, gap, G, threads)
    if rc != 0:
        raise ValueError(f"oracle encode_mt failed ({rc})")
    return pay[:W], gap[: (G + 7) // 8]


def decode_mt(payload, gap, payload_words, gap_elements, order_symbol, order_length, n,
              threads=0):
    pay = np.ascontiguousarray(payload, dtype=np.uint32)
    g = np.ascontiguousarray(gap, dtype=np.uint32)
    if pay.size == 0:
        pay = np.zeros(1, np.uint32)
    if g.size == 0:
        g = np.zeros(1, np.uint32)
    out = np.empty(max(n, 1), np.uint8)
    produced = load().orc_decode_segments_mt(
        pay, payload_words, g, gap_elements, np.ascontiguousarray(order_symbol, dtype=np.uint8),
        np.ascontiguousarray(order_length, dtype=np.uint8), len(order_symbol), out, n, threads)
    if produced < 0:
        raise ValueError(f"oracle decode_mt failed ({produced})")
    return out[:n]


def codebook(hist):
    """dict(k, code[256], length[256], order_symbol[k], order_length[k], flags)."""
    h = np.ascontiguousarray(hist, dtype=np.uint64)
    code = np.zeros(256, np.uint32)
    length = np.zeros(256, np.uint32)
    osym = np.zeros(256, np.uint8)
    olen = np.zeros(256, np.uint8)
    fl = ctypes.c_int(0)
    k = load().orc_build_codebook(h, code, length, osym, olen, ctypes.byref(fl))
    if k < 0:
        raise ValueError("oracle codebook failed")
    return dict(k=k, code=code, length=length, order_symbol=osym[:k].copy(),
                order_length=olen[:k].copy(), flags=fl.value)


def sizes(hist, length):
    out = [ctypes.c_uint64(0) for _ in range(4)]
    load().orc_sizes(np.ascontiguousarray(hist, dtype=np.uint64),
                     np.ascontiguousarray(length, dtype=np.uint32), *[ctypes.byref(o) for o in out])
    return dict(bits=out[0].value, payload_words=out[1].value, gap_elements=out[2].value,
                gap_words=out[3].value)


def encode(data, cb, sz=None):
    """(payload uint32[W], gap uint32[ceil(G/8)])."""
    d = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    if sz is None:
        sz = sizes(histogram(d), cb["length"])
    W, G = sz["payload_words"], sz["gap_elements"]
    pay = np.zeros(max(W, 1), np.uint32)
    gap = np.zeros(max((G + 7) // 8, 1), np.uint32)
    rc = load().orc_encode(d, d.size, cb["code"], cb["length"], pay, W, gap, G)
    if rc != 0:
        raise ValueError(f"oracle encode failed ({rc})")
    return pay[:W], gap[: (G + 7) // 8]


def decode(payload, gap, payload_words, gap_elements, order_symbol, order_length, n,
           seg_counts=False):
    pay = np.ascontiguousarray(payload, dtype=np.uint32)
    g = np.ascontiguousarray(gap, dtype=np.uint32)
    if pay.size == 0:
        pay = np.zeros(1, np.uint32)
    if g.size == 0:
        g = np.zeros(1, np.uint32)
    out = np.zeros(max(n, 1), np.uint8)
    counts = np.zeros(max(gap_elements, 1), np.uint32) if seg_counts else None
    produced = load().orc_decode_segments(
        pay, payload_words, g, gap_elements, np.ascontiguousarray(order_symbol, dtype=np.uint8),
        np.ascontiguousarray(order_length, dtype=np.uint8), len(order_symbol), out, n,
        counts.ctypes.data_as(ctypes.c_void_p) if seg_counts else None)
    if produced < 0:
        raise ValueError(f"oracle decode failed ({produced})")
    if seg_counts:
        return out[:n], int(produced), counts[:gap_elements]
    return out[:n]


def stream(cb, n, payload, gap, W, G):
    size = load().orc_stream_size(cb["k"], W, G)
    buf = np.zeros(size, np.uint8)
    pay = np.ascontiguousarray(payload, dtype=np.uint32)
    g = np.ascontiguousarray(gap, dtype=np.uint32)
    if pay.size == 0:
        pay = np.zeros(1, np.uint32)
    if g.size == 0:
        g = np.zeros(1, np.uint32)
    osym = np.zeros(256, np.uint8)
    olen = np.zeros(256, np.uint8)
    osym[: cb["k"]] = cb["order_symbol"]
    olen[: cb["k"]] = cb["order_length"]
    wrote = load().orc_stream_write(buf, size, cb["k"], osym, olen, n, W, G, g, pay)
    if wrote != size:
        raise ValueError("oracle stream write failed")
    return buf.tobytes()


def synth(n, start, seed, cdf, perm):
    out = np.zeros(max(n, 1), np.uint8)
    load().orc_synth_fill(out, n, start, seed, np.ascontiguousarray(cdf, dtype=np.uint64),
                          np.ascontiguousarray(perm, dtype=np.uint8))
    return out[:n]

"""Host-side mirror of the reference's encode/decode path on top of libhgap.so.

Device buffers are torch tensors on the current HIP device (torch is used for device memory
and streams only).  Every device call is asynchronous on `stream` (default: torch's current
stream) and goes through the C ABI in include/hgap.h.

Reference pipeline (paths relative to the reference repository root):
  histgram            encoder/src/encoder.cu:118-140      -> histogram()
  store_symbols+qsort encoder/src/symbols.cpp:29-43,
                      encoder/src/huff.cpp:17-21,78-79
  boundary_PM         encoder/src/package_merge.cpp:107   -> codebook_build()
  get_outputfilesize  encoder/src/symbols.cpp:45-55       -> encode_sizes()
  encode              encoder/src/encoder.cu:382-457      -> encode()
  stream writer       encoder/src/huff.cpp:111-125        -> stream_bytes() / compress()
  stream reader       decoder/src/huff.cpp:27-39,71-81    -> parse_stream() / decompress()
  get_twolevel_table  decoder/src/get_table.cpp:48-139    -> DeviceTable()
  decoder_l1_l2       decoder/src/decoder.cu:732-815      -> decode()
"""
import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _capi
from ._capi import Codebook, DectableInfo, HgapError, LaunchConfig, Sizes, check, default_config

__all__ = [
    "Codebook",
    "Sizes",
    "DectableInfo",
    "HgapError",
    "LaunchConfig",
    "default_config",
    "Encoded",
    "DeviceTable",
    "histogram",
    "codebook_build",
    "codebook_from_order",
    "encode_sizes",
    "encode",
    "decode",
    "stream_bytes",
    "parse_stream",
    "compress",
    "decompress",
    "synth_fill",
]


def _lib():
    return _capi.load()


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _as_device_u8(x):
    if isinstance(x, torch.Tensor):
        if x.device.type != "cuda":
            x = x.cuda()
        if x.dtype != torch.uint8:
            x = x.contiguous().view(torch.uint8)
        return x.contiguous().view(-1)
    arr = np.ascontiguousarray(np.frombuffer(x, dtype=np.uint8) if isinstance(x, (bytes, bytearray, memoryview)) else np.asarray(x, dtype=np.uint8)).reshape(-1)
    return torch.from_numpy(arr.copy()).cuda()


def _hist_array(hist):
    if isinstance(hist, torch.Tensor):
        hist = hist.detach().cpu().numpy()
    h = np.ascontiguousarray(np.asarray(hist).astype(np.uint32).reshape(-1))
    if h.size != 256:
        raise ValueError("a histogram has 256 bins")
    return h


def _words(n_words, device):
    return torch.empty(max(int(n_words), 4), dtype=torch.int32, device=device)


def histogram(d_in, stream=None):
    """256 byte counts of a device uint8 tensor (int32 tensor on the same device)."""
    d_in = _as_device_u8(d_in)
    hist = torch.empty(256, dtype=torch.int32, device=d_in.device)
    check(_lib().hgap_histogram(_ptr(d_in), d_in.numel(), _ptr(hist), _stream(stream)),
          "hgap_histogram")
    return hist


RANGE_BYTES = 256 * 1024  # HGAP_RANGE_BYTES
# HGAP_CHAINED_MIN_BYTES: below this the chained encoder has too few ranges to fill the device,
# and hgap_encode_ranges without a launch shape takes the per-tile look-back encoder
CHAINED_MIN_BYTES = 64 << 20


def histogram_ranges(d_in, stream=None, range_hist=None):
    """(hist, range_hist): the 256 byte counts, and the counts of every range of
    HGAP_RANGE_BYTES input bytes (int32 tensor of range_count(n) x 256) for encode()."""
    d_in = _as_device_u8(d_in)
    n = d_in.numel()
    hist = torch.empty(256, dtype=torch.int32, device=d_in.device)
    rows = int(_lib().hgap_range_count(n))
    if range_hist is None or range_hist.numel() < 256 * rows:
        range_hist = torch.empty((max(rows, 1), 256), dtype=torch.int32, device=d_in.device)
    check(_lib().hgap_histogram_ranges(_ptr(d_in), n, _ptr(hist), _ptr(range_hist),
                                       _stream(stream)), "hgap_histogram_ranges")
    return hist, range_hist


def host_histogram(hist):
    """Counts of a histogram tensor on the host as uint32 (the device tensor is int32: counts of
    2^31 and more read back as negative numbers unless reinterpreted)."""
    return hist.detach().cpu().numpy().view(np.uint32)


def codebook_build(hist):
    """Codebook from 256 counts, bit-exact with the reference's boundary_PM."""
    h = _hist_array(hist)
    cb = Codebook()
    check(_lib().hgap_codebook_build(h.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cb)),
          "hgap_codebook_build")
    return cb


def codebook_from_order(order_symbol, order_length):
    s = np.ascontiguousarray(order_symbol, dtype=np.uint8)
    l = np.ascontiguousarray(order_length, dtype=np.uint8)
    cb = Codebook()
    check(_lib().hgap_codebook_from_order(s.ctypes.data_as(ctypes.c_void_p),
                                          l.ctypes.data_as(ctypes.c_void_p), s.size,
                                          ctypes.byref(cb)), "hgap_codebook_from_order")
    return cb


def encode_sizes(cb, hist):
    h = _hist_array(hist)
    sz = Sizes()
    check(_lib().hgap_encode_sizes(ctypes.byref(cb), h.ctypes.data_as(ctypes.c_void_p),
                                   ctypes.byref(sz)), "hgap_encode_sizes")
    return sz


@dataclass
class Encoded:
    """Device-resident compressed buffer: payload[W] and gap[ceil(G/8)] as int32 words."""

    n: int
    codebook: Codebook
    sizes: Sizes
    payload: torch.Tensor
    gap: torch.Tensor

    @property
    def payload_words(self):
        return int(self.sizes.payload_words)

    @property
    def gap_elements(self):
        return int(self.sizes.gap_elements)

    def payload_u32(self):
        return self.payload[: self.payload_words].cpu().numpy().view(np.uint32)

    def gap_u32(self):
        return self.gap[: int(self.sizes.gap_words)].cpu().numpy().view(np.uint32)


class Workspace:
    """Reusable device workspace for the look-back status words (grows on demand)."""

    def __init__(self, device="cuda"):
        self.device = device
        self.buf = None

    def get(self, nbytes):
        nbytes = max(int(nbytes), 16)
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        return self.buf


def _cfg_ref(config):
    if config is None:
        return None
    if isinstance(config, dict):
        config = LaunchConfig.from_dict(config)
    return ctypes.byref(config)


def encode(d_in, cb=None, hist=None, sizes=None, stream=None, workspace=None, payload=None,
           gap=None, config=None, range_hist=None):
    """Encode a device uint8 buffer.  Without `cb`, builds the histogram and codebook first
    (device histogram, host codebook) like the reference's main (encoder/src/huff.cpp:69-99).
    With `range_hist` (from histogram_ranges over the same input) the encoder derives every
    range's bit offset from the counts and runs without waiting between tiles; the stream is
    the same."""
    d_in = _as_device_u8(d_in)
    n = d_in.numel()
    if cb is None or (sizes is None and hist is None):
        if hist is None:
            if n >= CHAINED_MIN_BYTES:
                hist, range_hist = histogram_ranges(d_in, stream)
            else:
                hist = histogram(d_in, stream)
        hist = _hist_array(hist)
        if cb is None:
            cb = codebook_build(hist)
    if sizes is None:
        sizes = encode_sizes(cb, hist)
    dev = d_in.device
    if payload is None:
        payload = _words(sizes.payload_words, dev)
    if gap is None:
        gap = _words(sizes.gap_words, dev)
    wsb = _lib().hgap_encode_workspace_bytes(n)
    ws = (workspace or Workspace(dev)).get(wsb)
    if range_hist is not None:
        check(_lib().hgap_encode_ranges(_ptr(d_in), n, ctypes.byref(cb), ctypes.byref(sizes),
                                        _ptr(range_hist), _ptr(payload), _ptr(gap), _ptr(ws),
                                        ws.numel(), _cfg_ref(config), _stream(stream)),
              "hgap_encode_ranges")
    elif config is None:  # default shape, falling back to smaller tiles when it does not fit
        check(_lib().hgap_encode(_ptr(d_in), n, ctypes.byref(cb), ctypes.byref(sizes),
                                 _ptr(payload), _ptr(gap), _ptr(ws), ws.numel(), _stream(stream)),
              "hgap_encode")
    else:
        check(_lib().hgap_encode_cfg(_ptr(d_in), n, ctypes.byref(cb), ctypes.byref(sizes),
                                     _ptr(payload), _ptr(gap), _ptr(ws), ws.numel(),
                                     _cfg_ref(config), _stream(stream)), "hgap_encode")
    return Encoded(n=n, codebook=cb, sizes=sizes, payload=payload, gap=gap)


class DeviceTable:
    """Decode table image for one codebook, resident in device memory."""

    def __init__(self, cb, prefix_bits=0, device="cuda"):
        info = DectableInfo()
        lib = _lib()
        check(lib.hgap_dectable_build(ctypes.byref(cb), prefix_bits, ctypes.byref(info), None, 0),
              "hgap_dectable_build")
        image = np.zeros(max(int(info.bytes) // 4, 4), dtype=np.uint32)
        if info.bytes:
            check(lib.hgap_dectable_build(ctypes.byref(cb), prefix_bits, ctypes.byref(info),
                                          image.ctypes.data_as(ctypes.c_void_p), image.nbytes),
                  "hgap_dectable_build")
        self.info = info
        self.image = image
        self.table = torch.from_numpy(image.view(np.int32)).to(device)

    @classmethod
    def resident(cls, info, table):
        """A table image already in device memory (`table`: int32 tensor, `info` its sizes)."""
        t = cls.__new__(cls)
        t.info = info
        t.image = None
        t.table = table
        return t


def decode(payload, gap, payload_words, gap_elements, table, n, out=None, stream=None,
           workspace=None, config=None):
    """Decode n bytes from device payload/gap words with a DeviceTable; returns uint8 tensor."""
    dev = payload.device
    if out is None:
        out = torch.empty(max(int(n), 16), dtype=torch.uint8, device=dev)
    wsb = _lib().hgap_decode_workspace_bytes(gap_elements)
    ws = (workspace or Workspace(dev)).get(wsb)
    if config is None:  # default shape, falling back to smaller tiles when it does not fit
        check(_lib().hgap_decode(_ptr(payload), payload_words, _ptr(gap), gap_elements,
                                 _ptr(table.table), ctypes.byref(table.info), _ptr(out), n,
                                 _ptr(ws), ws.numel(), _stream(stream)), "hgap_decode")
    else:
        check(_lib().hgap_decode_cfg(_ptr(payload), payload_words, _ptr(gap), gap_elements,
                                     _ptr(table.table), ctypes.byref(table.info), _ptr(out), n,
                                     _ptr(ws), ws.numel(), _cfg_ref(config), _stream(stream)),
              "hgap_decode")
    return out[:n]


def stream_bytes(enc):
    """The reference's compressed file for an Encoded buffer (encoder/src/huff.cpp:111-125)."""
    lib = _lib()
    k = enc.codebook.symbol_count
    hdr = np.zeros(int(lib.hgap_stream_header_bytes(k)), dtype=np.uint8)
    written = ctypes.c_size_t(0)
    check(lib.hgap_stream_write_header(ctypes.byref(enc.codebook), enc.n, ctypes.byref(enc.sizes),
                                       hdr.ctypes.data_as(ctypes.c_void_p), hdr.size,
                                       ctypes.byref(written)), "hgap_stream_write_header")
    return hdr.tobytes() + enc.gap_u32().astype("<u4").tobytes() + \
        enc.payload_u32().astype("<u4").tobytes()


def parse_stream(blob):
    """(codebook, n, sizes, gap uint32[], payload uint32[]) of a compressed file."""
    buf = np.frombuffer(blob, dtype=np.uint8)
    cb = Codebook()
    n = ctypes.c_uint32(0)
    sz = Sizes()
    goff = ctypes.c_uint64(0)
    poff = ctypes.c_uint64(0)
    check(_lib().hgap_stream_parse_header(buf.ctypes.data_as(ctypes.c_void_p), buf.size,
                                          ctypes.byref(cb), ctypes.byref(n), ctypes.byref(sz),
                                          ctypes.byref(goff), ctypes.byref(poff)),
          "hgap_stream_parse_header")
    gap = buf[goff.value: goff.value + 4 * sz.gap_words].view("<u4").astype(np.uint32)
    pay = buf[poff.value: poff.value + 4 * sz.payload_words].view("<u4").astype(np.uint32)
    return cb, int(n.value), sz, gap, pay


def compress(data, stream=None):
    """bytes/ndarray/tensor -> compressed file bytes (histogram and encode on the GPU)."""
    d_in = _as_device_u8(data)
    if d_in.numel() == 0:
        cb = Codebook()
        sz = Sizes()
        return stream_bytes(Encoded(0, cb, sz, _words(0, d_in.device), _words(0, d_in.device)))
    return stream_bytes(encode(d_in, stream=stream))


def decompress(blob, stream=None, device="cuda"):
    """compressed file bytes -> uint8 ndarray (decode on the GPU)."""
    cb, n, sz, gap, pay = parse_stream(blob)
    if n == 0:
        return np.zeros(0, dtype=np.uint8)
    table = DeviceTable(cb, device=device)
    d_pay = torch.from_numpy(np.concatenate([pay, np.zeros(4, np.uint32)]).view(np.int32)).to(device)
    d_gap = torch.from_numpy(np.concatenate([gap, np.zeros(4, np.uint32)]).view(np.int32)).to(device)
    out = decode(d_pay, d_gap, sz.payload_words, sz.gap_elements, table, n, stream=stream)
    return out.cpu().numpy()


def synth_fill(out, seed, cdf, perm, start=0, stream=None):
    """Fill a device uint8 tensor with the counter-based synthetic generator."""
    cdf = np.ascontiguousarray(cdf, dtype=np.uint64)
    perm = np.ascontiguousarray(perm, dtype=np.uint8)
    check(_lib().hgap_synth_fill(_ptr(out), out.numel(), start, seed,
                                 cdf.ctypes.data_as(ctypes.c_void_p),
                                 perm.ctypes.data_as(ctypes.c_void_p), _stream(stream)),
          "hgap_synth_fill")
    return out

"""ctypes view of libhgap.so (include/hgap.h, include/hgap_compat.h).

The shared library is built in-tree by ``make -C huffman_coding_gap_arrays_amd/csrc``
(``__graft_entry__.build()``).  Importing this module fails loudly when it is missing: there is
no CPU fallback for the device path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhgap.so")

HGAP_OK = 0
STATUS_NAMES = {
    0: "HGAP_OK",
    1: "HGAP_ERR_INVALID_ARG",
    2: "HGAP_ERR_SIZE",
    3: "HGAP_ERR_CODEBOOK",
    4: "HGAP_ERR_WORKSPACE",
    5: "HGAP_ERR_HIP",
    6: "HGAP_ERR_FORMAT",
    7: "HGAP_ERR_CAPACITY",
    8: "HGAP_ERR_UNSUPPORTED",
}


class HgapError(RuntimeError):
    """A libhgap entry point returned a non-zero hgap_status."""

    def __init__(self, status, where):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}")


class Codebook(ctypes.Structure):
    """hgap_codebook: Codetable[256] by byte value plus the stream's code-assignment order."""

    _fields_ = [
        ("symbol_count", ctypes.c_uint32),
        ("order_symbol", ctypes.c_uint8 * 256),
        ("order_length", ctypes.c_uint8 * 256),
        ("code", ctypes.c_uint32 * 256),
        ("length", ctypes.c_uint32 * 256),
    ]


class Sizes(ctypes.Structure):
    """hgap_sizes: bit count, payload words W, gap elements G, gap words ceil(G/8)."""

    _fields_ = [
        ("bits", ctypes.c_uint64),
        ("payload_words", ctypes.c_uint32),
        ("gap_elements", ctypes.c_uint32),
        ("gap_words", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class DectableInfo(ctypes.Structure):
    _fields_ = [
        ("prefix_bits", ctypes.c_uint32),
        ("l1_entries", ctypes.c_uint32),
        ("l2_entries", ctypes.c_uint32),
        ("bytes", ctypes.c_uint32),
        ("min_length", ctypes.c_uint32),
        ("max_length", ctypes.c_uint32),
        ("max_symbols_per_segment", ctypes.c_uint32),
        ("symbol_count", ctypes.c_uint32),
    ]


class LaunchConfig(ctypes.Structure):
    """hgap_launch_config: launch shape of the device steps."""

    _fields_ = [
        ("enc_threads", ctypes.c_uint32),
        ("enc_bytes_per_thread", ctypes.c_uint32),
        ("dec_threads", ctypes.c_uint32),
        ("lookback_window", ctypes.c_uint32),
        ("blocks_per_cu", ctypes.c_uint32),
        ("prefix_bits", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}

    @classmethod
    def from_dict(cls, d):
        c = default_config()
        for k, v in d.items():
            setattr(c, k, int(v))
        return c


def default_config():
    c = LaunchConfig()
    load().hgap_default_config(ctypes.byref(c))
    return c


def default_chained_config():
    c = LaunchConfig()
    load().hgap_default_chained_config(ctypes.byref(c))
    return c


class HistJob(ctypes.Structure):
    """hgap_hist_job (one buffer of a batched histogram)."""

    _fields_ = [("in_", ctypes.c_uint64), ("n", ctypes.c_uint64), ("hist", ctypes.c_uint64),
                ("range_base", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("base", ctypes.c_uint64)]


class EncJob(ctypes.Structure):
    """hgap_enc_job (one buffer of a batched encode)."""

    _fields_ = [("in_", ctypes.c_uint64), ("n", ctypes.c_uint64), ("payload", ctypes.c_uint64),
                ("gap", ctypes.c_uint64), ("payload_words", ctypes.c_uint32),
                ("gap_words", ctypes.c_uint32), ("table", ctypes.c_uint32),
                ("max_len", ctypes.c_uint32), ("base", ctypes.c_uint32),
                ("num_tiles", ctypes.c_uint32), ("range_base", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class DecJob(ctypes.Structure):
    """hgap_dec_job (one buffer of a batched decode)."""

    _fields_ = [("payload", ctypes.c_uint64), ("gap", ctypes.c_uint64), ("out", ctypes.c_uint64),
                ("n", ctypes.c_uint64), ("table", ctypes.c_uint64),
                ("payload_words", ctypes.c_uint32), ("gap_elements", ctypes.c_uint32),
                ("prefix_bits", ctypes.c_uint32), ("l1_entries", ctypes.c_uint32),
                ("table_words", ctypes.c_uint32), ("max_symbols", ctypes.c_uint32),
                ("max_length", ctypes.c_uint32), ("base", ctypes.c_uint32),
                ("num_tiles", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class Symbol(ctypes.Structure):
    """struct Symbol (encoder/include/constants.hpp:34-38)."""

    _fields_ = [("symbol", ctypes.c_uint8), ("length", ctypes.c_uint8), ("num", ctypes.c_uint32)]


class Codetable(ctypes.Structure):
    """struct Codetable (encoder/include/constants.hpp:40-43)."""

    _fields_ = [("code", ctypes.c_uint32), ("length", ctypes.c_uint32)]


class TableInfo(ctypes.Structure):
    """struct TableInfo (decoder/include/constants.hpp:53-57)."""

    _fields_ = [
        ("l1table_size", ctypes.c_uint32),
        ("l2table_size", ctypes.c_uint32),
        ("ptrtable_size", ctypes.c_uint32),
    ]


_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_SZ = ctypes.c_size_t
_INT = ctypes.c_int

# name -> (restype, argtypes); every function declared in include/hgap.h and hgap_compat.h
PROTOTYPES = {
    "hgap_status_string": (ctypes.c_char_p, [_INT]),
    "hgap_device_cu_count": (_INT, []),
    "hgap_default_config": (None, [ctypes.c_void_p]),
    "hgap_default_chained_config": (None, [ctypes.c_void_p]),
    "hgap_codebook_build": (_INT, [_P, ctypes.POINTER(Codebook)]),
    "hgap_codebook_from_order": (_INT, [_P, _P, _U32, ctypes.POINTER(Codebook)]),
    "hgap_encode_sizes": (_INT, [ctypes.POINTER(Codebook), _P, ctypes.POINTER(Sizes)]),
    "hgap_dectable_build": (
        _INT,
        [ctypes.POINTER(Codebook), _U32, ctypes.POINTER(DectableInfo), _P, _SZ],
    ),
    "hgap_stream_bytes": (_U64, [_U32, _U32, _U32]),
    "hgap_stream_header_bytes": (_U64, [_U32]),
    "hgap_stream_write_header": (
        _INT,
        [ctypes.POINTER(Codebook), _U32, ctypes.POINTER(Sizes), _P, _SZ, ctypes.POINTER(_SZ)],
    ),
    "hgap_stream_parse_header": (
        _INT,
        [
            _P,
            _SZ,
            ctypes.POINTER(Codebook),
            ctypes.POINTER(_U32),
            ctypes.POINTER(Sizes),
            ctypes.POINTER(_U64),
            ctypes.POINTER(_U64),
        ],
    ),
    "hgap_histogram": (_INT, [_P, _U64, _P, _P]),
    "hgap_encode_workspace_bytes": (_SZ, [_U64]),
    "hgap_encode": (
        _INT,
        [_P, _U64, ctypes.POINTER(Codebook), ctypes.POINTER(Sizes), _P, _P, _P, _SZ, _P],
    ),
    "hgap_encode_cfg": (
        _INT,
        [_P, _U64, ctypes.POINTER(Codebook), ctypes.POINTER(Sizes), _P, _P, _P, _SZ,
         ctypes.c_void_p, _P],
    ),
    "hgap_range_count": (_U32, [_U64]),
    "hgap_histogram_ranges": (_INT, [_P, _U64, _P, _P, _P]),
    "hgap_encode_ranges": (
        _INT,
        [_P, _U64, ctypes.POINTER(Codebook), ctypes.POINTER(Sizes), _P, _P, _P, _P, _SZ,
         ctypes.c_void_p, _P],
    ),
    "hgap_decode_workspace_bytes": (_SZ, [_U32]),
    "hgap_decode_cfg": (
        _INT,
        [_P, _U32, _P, _U32, _P, ctypes.POINTER(DectableInfo), _P, _U64, _P, _SZ,
         ctypes.c_void_p, _P],
    ),
    "hgap_decode": (
        _INT,
        [_P, _U32, _P, _U32, _P, ctypes.POINTER(DectableInfo), _P, _U64, _P, _SZ, _P],
    ),
    "hgap_batch_plan_histogram": (_U64, [_P, _U32]),
    "hgap_batch_histogram": (_INT, [_P, _U32, _U64, _P]),
    "hgap_codebook_build_batch": (_INT, [_P, _U32, _P, _U32]),
    "hgap_encode_table": (_INT, [ctypes.POINTER(Codebook), _P]),
    "hgap_batch_plan_encode": (_U32, [_P, _U32, ctypes.c_void_p]),
    "hgap_batch_plan_decode": (_U32, [_P, _U32, ctypes.c_void_p]),
    "hgap_batch_workspace_bytes": (_SZ, [_U32]),
    "hgap_batch_decode_workspace_bytes": (_SZ, [_U32, ctypes.POINTER(LaunchConfig)]),
    "hgap_batch_encode": (_INT, [_P, _U32, _U32, _U32, _P, _P, _SZ, ctypes.c_void_p, _P]),
    "hgap_batch_histogram_ranges": (_INT, [_P, _U32, _U32, _P, _P]),
    "hgap_batch_plan_encode_ranges": (_U32, [_P, _U32, ctypes.c_void_p]),
    "hgap_batch_encode_ranges_workspace_bytes": (_SZ, [_U32]),
    "hgap_batch_encode_ranges": (_INT, [_P, _U32, _U32, _U32, _P, _P, _P, _SZ, ctypes.c_void_p,
                                        _P]),
    "hgap_batch_decode": (_INT, [_P, _U32, _U32, _U32, _U32, _P, _SZ, ctypes.c_void_p, _P]),
    "hgap_synth_fill": (_INT, [_P, _U64, _U64, _U64, _P, _P, _P]),
    "hgap_batch_prepare": (_INT, [_P, _U32, _P, _P, _P, _U32]),
    "hgap_dectable_build_batch": (_INT, [_P, _U32, _P, _P, _P, _P, _SZ, _U32]),
    "hgap_config_supported": (_INT, [ctypes.c_void_p, _U32, ctypes.POINTER(DectableInfo)]),
    "hgap_decode_image_bytes": (_U32, [_U32, _U32, _U32]),
    "hgap_mlp_param_count": (_U32, [_U32, _U32, _U32]),
    "hgap_mlp_forward": (_INT, [_P, _U32, _U32, _U32, _U32, _P, _P, _P, _P]),
    "hgap_mlp_backward": (
        _INT,
        [_P, _P, _P, _P, _P, _U32, _U32, _U32, _U32, ctypes.c_float, _P, _P, _P, _P, _P],
    ),
    "hgap_mlp_adam_step": (
        _INT,
        [_P, _P, _P, _P, _U32, _U32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
         ctypes.c_float, _P],
    ),
    # reference-named entry points (include/hgap_compat.h)
    "histgram": (None, [_P, _P, _U32]),
    "encode": (None, [_P, _U32, _P, _U32, _U32, _P]),
    "store_symbols": (_INT, [_P, _P]),
    "get_outputfilesize": (ctypes.c_ulonglong, [_P, _U32]),
    "boundary_PM": (None, [_P, _U32, _P]),
    "get_table_info": (_U32, [_P, _INT, _INT, ctypes.POINTER(TableInfo)]),
    "get_twolevel_table": (_U32, [_P, _INT, _P, _INT, TableInfo]),
    "decoder_l1_l2": (None, [_P, _INT, _P, _INT, _INT, _P, _INT, _U32, _U32, TableInfo]),
}

_lib = None


def load():
    """Load libhgap.so (once) and attach the prototypes; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
            "g.build()'` (make -C huffman_coding_gap_arrays_amd/csrc)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, where):
    if status != HGAP_OK:
        raise HgapError(status, where)

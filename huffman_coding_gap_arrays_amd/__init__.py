"""MI355X-native Huffman coding with gap arrays.

A drop-in for the encode/decode path of daisuke-takafuji/Huffman_coding_Gap_arrays: the same
compressed-stream layout (codebook header, MSB-first 32-bit payload, 4-bit-per-128-bit-segment
gap array), produced and consumed by hand-written gfx950 kernels in libhgap.so behind the C ABI
of include/hgap.h.  The reference's own entry points (histgram, encode, boundary_PM,
get_table_info, get_twolevel_table, decoder_l1_l2, ...) are exported under their names too
(include/hgap_compat.h).
"""
from . import synth  # noqa: F401  (numpy only)
from ._capi import HgapError, load  # noqa: F401

__all__ = ["synth", "HgapError", "load", "codec"]


def __getattr__(name):
    # torch-backed device API, imported lazily so that CPU-only tools need not import torch
    if name == "codec":
        import importlib

        return importlib.import_module(__name__ + ".codec")
    raise AttributeError(name)

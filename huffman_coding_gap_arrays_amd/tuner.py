"""Exhaustive autotuner for the launch shape of the encoder and the decoder.

The searchable knobs are the fields of hgap_launch_config (include/hgap.h) plus the decode
table's first-level width:
  encoder: workgroup size x input bytes per thread (tile = their product), look-back window
  decoder: workgroup size (segments per tile), look-back window, first-level table bits
Every candidate is timed with HIP events on the stream the kernels run on (median of `reps`
launches after `warmup`), on the caller's buffer.  Shapes the device cannot hold (LDS) are
skipped.  The best shape per buffer is the reference label the learned predictor is trained
on, and the judge of its regret.

usage: python -m huffman_coding_gap_arrays_amd.tuner [--config zipf_1GiB] [--reps 5]
"""
import argparse
import itertools
import json
import statistics
import sys

import torch

from . import codec
from ._capi import HgapError, LaunchConfig, default_config

ENC_SHAPES = [(t, k) for t in (256, 512, 1024) for k in (16, 32, 64)]
DEC_THREADS = (256, 512, 1024)
PREFIX_BITS = (8, 9, 10, 11, 12)
WINDOWS = (64, 128, 256)


def enc_space(windows=WINDOWS):
    for (t, k), w in itertools.product(ENC_SHAPES, windows):
        yield dict(enc_threads=t, enc_bytes_per_thread=k, lookback_window=w)


def dec_space(windows=WINDOWS, prefixes=PREFIX_BITS):
    for t, w, p in itertools.product(DEC_THREADS, windows, prefixes):
        yield dict(dec_threads=t, lookback_window=w, prefix_bits=p)


def _median_ms(fn, reps, warmup, stream):
    for _ in range(warmup):
        fn()
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    return statistics.median(times)


def time_encode(d_in, cb, sizes, config, reps=5, warmup=1, workspace=None, payload=None,
                gap=None, stream=None):
    """Median ms of one hgap_encode_cfg launch, or None when the shape is unsupported."""
    stream = stream or torch.cuda.current_stream()
    ws = workspace or codec.Workspace(d_in.device)
    cfg = LaunchConfig.from_dict(config)
    enc = None

    def run():
        nonlocal enc
        enc = codec.encode(d_in, cb=cb, sizes=sizes, stream=stream, workspace=ws,
                           payload=payload, gap=gap, config=cfg)

    try:
        return _median_ms(run, reps, warmup, stream)
    except HgapError as e:
        if e.status == 8:
            return None
        raise


def time_decode(enc, table, config, out=None, reps=5, warmup=1, workspace=None, stream=None):
    stream = stream or torch.cuda.current_stream()
    ws = workspace or codec.Workspace(enc.payload.device)
    cfg = LaunchConfig.from_dict(config)

    def run():
        codec.decode(enc.payload, enc.gap, enc.payload_words, enc.gap_elements, table, enc.n,
                     out=out, stream=stream, workspace=ws, config=cfg)

    try:
        return _median_ms(run, reps, warmup, stream)
    except HgapError as e:
        if e.status == 8:
            return None
        raise


def tune_encode(d_in, cb, sizes, space=None, reps=5, warmup=1):
    """(best config dict, [(config, ms or None), ...]) over the encoder search space."""
    ws = codec.Workspace(d_in.device)
    payload = codec._words(sizes.payload_words, d_in.device)
    gap = codec._words(sizes.gap_words, d_in.device)
    results = []
    for cfg in (space if space is not None else enc_space()):
        ms = time_encode(d_in, cb, sizes, cfg, reps, warmup, ws, payload, gap)
        results.append((cfg, ms))
    ok = [r for r in results if r[1] is not None]
    return min(ok, key=lambda r: r[1])[0], results


def tune_decode(enc, space=None, reps=5, warmup=1):
    """(best config dict, [(config, ms or None), ...]) over the decoder search space."""
    ws = codec.Workspace(enc.payload.device)
    out = torch.empty(enc.n + 16, dtype=torch.uint8, device=enc.payload.device)
    tables = {}
    results = []
    for cfg in (space if space is not None else dec_space()):
        p = cfg.get("prefix_bits", 0)
        if p not in tables:
            tables[p] = codec.DeviceTable(enc.codebook, prefix_bits=p, device=enc.payload.device)
        ms = time_decode(enc, tables[p], cfg, out, reps, warmup, ws)
        results.append((cfg, ms))
    ok = [r for r in results if r[1] is not None]
    return min(ok, key=lambda r: r[1])[0], results


def main():
    from . import synth

    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="zipf_1GiB")
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    n, seed, cdf, perm = synth.spec(args.config)
    d_in = torch.empty(n, dtype=torch.uint8, device="cuda")
    codec.synth_fill(d_in, seed, cdf, perm)
    hist = codec.host_histogram(codec.histogram(d_in))
    cb = codec.codebook_build(hist)
    sizes = codec.encode_sizes(cb, hist)
    C = 4 * (sizes.payload_words + sizes.gap_words)
    best_e, res_e = tune_encode(d_in, cb, sizes, reps=args.reps)
    for cfg, ms in res_e:
        print(json.dumps(dict(stage="encode", **cfg, ms=ms,
                              alg_GBps=None if ms is None else round((n + C) / ms / 1e6, 1))),
              flush=True)
    enc = codec.encode(d_in, cb=cb, sizes=sizes, config=best_e)
    best_d, res_d = tune_decode(enc, reps=args.reps)
    for cfg, ms in res_d:
        print(json.dumps(dict(stage="decode", **cfg, ms=ms,
                              alg_GBps=None if ms is None else round((n + C) / ms / 1e6, 1))),
              flush=True)
    print(json.dumps(dict(best_encode=best_e, best_decode=best_d,
                          default=default_config().as_dict())), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

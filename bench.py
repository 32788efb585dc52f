"""Benchmark of the Huffman-with-gap-arrays hot path on one MI355X.

One step = one pass of the path over the 1 GiB Zipf(s=1.1) buffer (uint8, 256 ranks, seeded
rank-to-byte permutation), already resident in HBM:
    histogram (GPU) -> codebook (host, boundary package-merge) -> encode (GPU) -> decode (GPU)
`value` is uncompressed GB/s of the whole step.  Per-stage GB/s and the roofline of the
dominant kernel come from HIP events recorded on the stream the kernels run on.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config zipf_1GiB]
Prints exactly one JSON line on stdout.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(d_dev, sample_bytes, repeats, threads):
    """The oracle (C restatement of the reference path) on the first `sample_bytes` of the same
    buffer: histogram + codebook + encode + decode per repeat, on `threads` cores (the OpenMP
    variants) and on one core (the serial functions)."""
    from oracle import oracle

    oracle.load()
    sample = d_dev[:sample_bytes].cpu().numpy()

    def run(mt):
        t0 = time.perf_counter()
        h = oracle.histogram_mt(sample, threads) if mt else oracle.histogram(sample)
        cb = oracle.codebook(h)
        sz = oracle.sizes(h, cb["length"])
        if mt:
            pay, gap = oracle.encode_mt(sample, cb, sz, threads)
            out = oracle.decode_mt(pay, gap, sz["payload_words"], sz["gap_elements"],
                                   cb["order_symbol"], cb["order_length"], sample.size, threads)
        else:
            pay, gap = oracle.encode(sample, cb, sz)
            out = oracle.decode(pay, gap, sz["payload_words"], sz["gap_elements"],
                                cb["order_symbol"], cb["order_length"], sample.size)
        t = time.perf_counter() - t0
        if not (out == sample).all():
            raise RuntimeError("oracle round trip failed")
        return t

    run(True)  # warm-up (page faults of the output buffers, thread pool start)
    t_mt = statistics.median(run(True) for _ in range(repeats))
    t_1 = run(False)
    what = "histogram+codebook+encode+decode"
    return dict(value=round(sample.size / t_mt / 1e9, 4), unit="GB/s", cores=threads,
                kind="port",
                sample=f"first {sample.size >> 20} MiB of the same buffer; {what} with the "
                       f"oracle's OpenMP variants on {threads} threads, median of {repeats}",
                seconds=round(t_mt, 4),
                single_thread=dict(value=round(sample.size / t_1 / 1e9, 4), unit="GB/s",
                                   cores=1, seconds=round(t_1, 3)),
                host_cpus=os.cpu_count())


class _Hip:
    """hipMalloc / hipMemcpy through ctypes (the reference-named encode() frees its input with
    hipFree, so that input cannot be a torch allocation)."""

    def __init__(self):
        import ctypes

        self.c = ctypes
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]

    def malloc(self, n):
        p = self.c.c_void_p()
        if self.lib.hipMalloc(self.c.byref(p), n) != 0:
            raise RuntimeError("hipMalloc failed")
        return p


def compat_pcie(d_in, n):
    """The reference-named entry points, which take host buffers, on the same buffer:
    encode() (device input -> host payload + gap, encoder.cu:382-457 semantics) and
    decoder_l1_l2() (host stream -> host bytes, decoder.cu:732-815 semantics).  Their rates
    include the PCIe copies and the per-call allocations."""
    import ctypes

    from huffman_coding_gap_arrays_amd import _capi, codec
    from huffman_coding_gap_arrays_amd._capi import Codetable, Symbol, TableInfo

    lib = _capi.load()
    hip = _Hip()
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    hist = codec.host_histogram(codec.histogram(d_in))
    syms = (Symbol * 256)()
    k = lib.store_symbols(vp(hist), syms)
    srt = (Symbol * 256)()
    for j, o in enumerate(sorted(range(k), key=lambda j: (syms[j].num, syms[j].symbol))):
        srt[j] = syms[o]
    table = (Codetable * 256)()
    lib.boundary_PM(srt, k, table)
    bits = lib.get_outputfilesize(srt, k)
    G, W = (bits + 127) // 128, (bits + 31) // 32
    host_out = np.zeros(W + (G + 7) // 8, np.uint32)
    words = (n + 3) // 4
    times = []
    for _ in range(2):  # the first call warms up the host buffers
        d_copy = hip.malloc(4 * (words + 1))
        hip.lib.hipMemcpy(d_copy, ctypes.c_void_p(d_in.data_ptr()), n, 3)
        t0 = time.perf_counter()
        lib.encode(vp(host_out), W, d_copy, n, G, table)  # frees d_copy
        times.append(time.perf_counter() - t0)
    t_enc = times[-1]
    fsyms = (Symbol * 256)()
    for j in range(k):
        fsyms[j] = srt[k - 1 - j]
    info = TableInfo()
    P = lib.get_table_info(fsyms, k, 10, ctypes.byref(info))
    size = 4 * info.ptrtable_size + 256 + info.l1table_size + info.l2table_size
    stage = np.zeros(size, np.uint8)
    lib.get_twolevel_table(vp(stage), P, fsyms, k, info)
    inp = np.concatenate([host_out[W:], host_out[:W]])
    res = np.zeros(4 * words, np.uint8)
    # the reference decoder prints its own timing line on stdout: keep stdout for the JSON line
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        times = []
        for _ in range(2):
            t0 = time.perf_counter()
            lib.decoder_l1_l2(vp(inp), W, vp(res), n, G, vp(stage), size, P, k, info)
            times.append(time.perf_counter() - t0)
    finally:
        os.dup2(saved, 1)
        os.close(saved)
    t_dec = times[-1]
    ok = bool((res[:n] == d_in.cpu().numpy()).all())
    return dict(encode_GBps=round(n / t_enc / 1e9, 2), encode_s=round(t_enc, 4),
                decode_GBps=round(n / t_dec / 1e9, 2), decode_s=round(t_dec, 4),
                round_trip_exact=ok,
                note="reference-named entry points with host buffers: PCIe copies and per-call "
                     "allocations included (uncompressed bytes per second)")


def load_traffic(config, kernel):
    """HBM bytes per launch of `kernel` (or per decode step: "decode") measured by rocprofv3
    --pmc (profiles/*.json), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{config}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="zipf_1GiB")
    ap.add_argument("--cpu-sample-mib", type=int, default=256)
    ap.add_argument("--cpu-repeats", type=int, default=5)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="threads of the CPU baseline (the GPU box's CPU share is 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compat", action="store_true",
                    help="also time the reference-named host-buffer entry points (PCIe included)")
    ap.add_argument("--batch-count", type=int, default=4096,
                    help="buffers of the batched workload (0 = skip it)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py measures one GPU only: this build has no multi-GPU path "
                 "(the encoder would shard by input range with a 256-bin all-reduce and a scalar "
                 "exclusive scan; see DESIGN.md).  Run with --gpus 1.")

    import torch

    from huffman_coding_gap_arrays_amd import codec, synth

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    n, seed, cdf, perm = synth.spec(args.config)
    d_in = torch.empty(n, dtype=torch.uint8, device=dev)
    codec.synth_fill(d_in, seed, cdf, perm)

    # sizes/buffers from one untimed pass
    hist_dev, range_hist = codec.histogram_ranges(d_in)
    hist = codec.host_histogram(hist_dev)
    cb = codec.codebook_build(hist)
    sizes = codec.encode_sizes(cb, hist)
    W, G, GW = int(sizes.payload_words), int(sizes.gap_elements), int(sizes.gap_words)
    payload = torch.empty(W + 4, dtype=torch.int32, device=dev)
    gap = torch.empty(GW + 4, dtype=torch.int32, device=dev)
    out = torch.empty(n + 16, dtype=torch.uint8, device=dev)
    ws = codec.Workspace(dev)
    ws.get(max(codec._lib().hgap_encode_workspace_bytes(n),
               codec._lib().hgap_decode_workspace_bytes(G)))
    C = 4 * (W + GW)  # compressed bytes (payload + gap)

    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731

    def step(record):
        e = [ev() for _ in range(5)] if record else None
        if record:
            e[0].record(stream)
        h, rh = codec.histogram_ranges(d_in, stream, range_hist=range_hist)
        if record:
            e[1].record(stream)
        hb = codec.host_histogram(h)  # D2H of 256 counts (synchronizes), as the reference does
        book = codec.codebook_build(hb)
        sz = codec.encode_sizes(book, hb)
        table = codec.DeviceTable(book, device=dev)
        if record:
            e[2].record(stream)
        enc = codec.encode(d_in, cb=book, sizes=sz, stream=stream, workspace=ws,
                           payload=payload, gap=gap, range_hist=rh)
        if record:
            e[3].record(stream)
        codec.decode(enc.payload, enc.gap, enc.payload_words, enc.gap_elements, table, n,
                     out=out, stream=stream, workspace=ws)
        if record:
            e[4].record(stream)
        return e

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    recs = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        recs.append(step(True))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_hist = statistics.mean(r[0].elapsed_time(r[1]) for r in recs) / 1e3
    t_book = statistics.mean(r[1].elapsed_time(r[2]) for r in recs) / 1e3
    t_enc = statistics.mean(r[2].elapsed_time(r[3]) for r in recs) / 1e3
    t_dec = statistics.mean(r[3].elapsed_time(r[4]) for r in recs) / 1e3
    ms_per_step = wall / args.steps * 1e3

    ok = bool(torch.equal(out[:n], d_in))
    if not ok:
        sys.exit("decoded output differs from the input")

    stages = {
        "histogram": dict(seconds=t_hist, alg_bytes=n),
        "encode": dict(seconds=t_enc, alg_bytes=n + C),
        "decode": dict(seconds=t_dec, alg_bytes=C + n),
    }
    per_stage = {}
    for k, v in stages.items():
        gbs = v["alg_bytes"] / v["seconds"] / 1e9
        per_stage[k] = dict(ms=round(v["seconds"] * 1e3, 4),
                            uncompressed_GBps=round(n / v["seconds"] / 1e9, 2),
                            alg_GBps=round(gbs, 2), roofline_frac=round(gbs / HBM_PEAK_GBS, 4))
    per_stage["codebook_host"] = dict(ms=round(t_book * 1e3, 4))
    # the slower of encode (range prep and encoder kernels) and decode (count, offsets and write
    # kernels after a workspace fill), each timed as one unit between events on their stream
    dom = max(("encode", "decode"), key=lambda k: stages[k]["seconds"])
    kern = {"encode": "enc_range_prep_kernel+enc_kernel",
            "decode": "dec_count_kernel+dec_offsets_kernel+dec_write_kernel"}[dom]
    achieved = stages[dom]["alg_bytes"] / stages[dom]["seconds"] / 1e9
    roofline = dict(bound="hbm", kernel=kern, achieved=round(achieved, 2), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=load_traffic(args.config, dom),
                    alg_bytes_per_launch=stages[dom]["alg_bytes"])
    batched = None
    if args.batch_count > 0:
        log("batched workload ...")
        del out
        from huffman_coding_gap_arrays_amd import batch, predictor

        variants = [("default", None, None)]
        if os.path.exists(predictor.DEFAULT_MODEL):
            variants.append(("learned", None, predictor.Predictor.load(predictor.DEFAULT_MODEL)))
        batched = batch.bench(count=args.batch_count, variants=variants)
        batched["workload"] = (f"{args.batch_count} buffers, sizes log-uniform 4 KiB..64 MiB, "
                               "Zipf s ~ U(0.6, 2.0), own permutation and codebook each")
    compat = None
    if args.compat:
        log("reference-named host-buffer entry points ...")
        compat = compat_pcie(d_in, n)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(d_in, min(n, args.cpu_sample_mib << 20), args.cpu_repeats,
                           args.cpu_threads)
    value = n / (ms_per_step / 1e3) / 1e9
    line = {
        "metric": "encode/decode GB/s vs 8 TB/s HBM roofline (uncompressed bytes per second of "
                  "histogram+codebook+encode+decode)",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (counter-based splitmix64 generator on the GPU)",
        "config": {"workload": f"{args.config}: {n} bytes uint8, "
                               + ("Zipf s=1.1 over 256 ranks, seeded rank-to-byte permutation"
                                  if args.config == "zipf_1GiB" else args.config),
                   "n_bytes": n, "compressed_bytes": C, "payload_words": W, "gap_elements": G,
                   "symbols": int(cb.symbol_count), "parallelism": "single GPU"},
        "stages": per_stage,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "compat_pcie": compat,
        "round_trip_exact": ok,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Request loop serving mixed encode and decode requests with continuous batching.

Requests arrive over time (Poisson arrivals in the benchmark).  The loop admits queued requests
in arrival order while the device-memory budget `cap_bytes` and the in-flight batch slots allow,
turns whatever has been admitted into one batch, and launches it on the next of several streams,
so new batches overlap on the device with those still running.  A batch carries both kinds:
  decode requests -> one batched decode launch per launch-shape group, right away;
  encode requests -> one batched histogram launch; when it completes, threaded host codebooks,
                     exact sizes (the reservation shrinks from its worst-case bound), then one
                     batched encode launch per shape group.
Decode tables are cached on the device by codebook, so a codebook seen before is not rebuilt or
uploaded again.  A request whose worst-case memory exceeds the budget is rejected.

Latency = time the host observed the request's batch complete - the request's arrival time.
Throughput = uncompressed bytes of all served requests / (last completion - first arrival).

usage: python -m huffman_coding_gap_arrays_amd.serve [--requests 2000] [--load 0.65] [--rate R]
"""
import argparse
import ctypes
import json
import math
import statistics
import sys
import time
from collections import OrderedDict, deque
from dataclasses import dataclass

import numpy as np
import torch

from . import _capi, batch, codec
from ._capi import LaunchConfig, Sizes, check

ALIGN = 256
GB = 1e9


def _up(b, x):
    return (x + b - 1) // b * b


def _vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _sp(stream):
    return ctypes.c_void_p(stream.cuda_stream)


@dataclass
class Request:
    """One request.  encode: `data` holds the raw bytes (device uint8).  decode: `compressed`
    holds the payload/gap words and codebook of a stream (codec.Encoded, device)."""

    rid: int
    kind: str
    n: int
    arrival: float = 0.0
    data: torch.Tensor = None
    compressed: codec.Encoded = None
    # response
    result: object = None  # encode: codec.Encoded; decode: uint8 tensor [n]
    admitted: float = None
    done: float = None
    error: str = None
    reserved: int = 0
    batch_size: int = 0

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


def _range_bytes(n):
    """Per-range counts (1 KiB per 256 KiB range) and the chained encoder's status word and job
    index per range."""
    return 1036 * (-(-n // codec.RANGE_BYTES))


def encode_reserve_bound(n):
    """Device bytes an encode request may need before its histogram is known: payload for codes
    of at most 16 bits, the gap array, look-back status words for the smallest tile, its
    histogram and range counts, and allocation rounding."""
    words = (16 * n + 31) // 32
    gap_words = ((16 * n + 127) // 128 + 7) // 8
    return (4 * words + 4 * gap_words + 8 * (n // 4096 + 2) + _range_bytes(n) + 1024
            + 4 * ALIGN)


def encode_reserve_exact(n, sizes):
    return (4 * int(sizes.payload_words) + 4 * int(sizes.gap_words) + 8 * (n // 4096 + 2)
            + _range_bytes(n) + 1024 + 4 * ALIGN)


def decode_reserve(n, gap_elements):
    """Output, plus the request's share of the batched decoder's workspace: one count byte and
    1/16 byte of wave sums per segment (whole 1024-segment tiles), one offset per 256-segment
    tile, the offsets scan's status words."""
    segs = (gap_elements + 1023) // 1024 * 1024
    return _up(ALIGN, n + 16) + segs + segs // 16 + 8 * (gap_elements // 256 + 2) + 64 + 4 * ALIGN


def request_bytes(r):
    """Device bytes reserved when a request is admitted."""
    if r.kind == "encode":
        return encode_reserve_bound(r.n)
    return decode_reserve(r.n, r.compressed.gap_elements)


class TableCache:
    """Decode tables resident in device memory, keyed by codebook (LRU within `cap_bytes`)."""

    def __init__(self, device, cap_bytes=256 << 20):
        self.device = device
        self.cap = cap_bytes
        self.bytes = 0
        self.entries = OrderedDict()
        self.hits = 0
        self.misses = 0

    @staticmethod
    def key(cb):
        k = int(cb.symbol_count)
        return bytes(cb.order_symbol[:k]) + bytes(cb.order_length[:k])

    def get(self, cb, prefix_bits=0):
        key = (self.key(cb), prefix_bits)
        t = self.entries.get(key)
        if t is not None:
            self.hits += 1
            self.entries.move_to_end(key)
            return t
        self.misses += 1
        return self._put(key, codec.DeviceTable(cb, prefix_bits, self.device))

    def _put(self, key, t):
        self.entries[key] = t
        self.bytes += t.table.numel() * 4
        while self.bytes > self.cap and len(self.entries) > 1:
            _, old = self.entries.popitem(last=False)
            self.bytes -= old.table.numel() * 4
        return t

    def get_many(self, cbs, prefix_bits):
        """Tables for several codebooks.  Hits come from the cache; the missing tables are built
        together on the host (hgap_dectable_build_batch) and reach the device in one copy
        (synchronous, so that a batch on another stream can use them at once)."""
        keys = [(self.key(cb), int(p)) for cb, p in zip(cbs, prefix_bits)]
        out = [None] * len(keys)
        miss = OrderedDict()
        for i, k in enumerate(keys):
            t = self.entries.get(k)
            if t is not None:
                self.hits += 1
                self.entries.move_to_end(k)
                out[i] = t
            elif k in miss:
                self.hits += 1  # the same codebook twice in one call: built once
                miss[k].append(i)
            else:
                self.misses += 1
                miss[k] = [i]
        if not miss:
            return out
        lib = _capi.load()
        ks = list(miss)
        m = len(ks)
        arr = (_capi.Codebook * m)()
        for j, k in enumerate(ks):
            arr[j] = cbs[miss[k][0]]
        pb = np.array([k[1] for k in ks], np.uint32)
        infos = (_capi.DectableInfo * m)()
        check(lib.hgap_dectable_build_batch(arr, m, _vp(pb), infos, None, None, 0, 1),
              "hgap_dectable_build_batch")
        sizes = np.frombuffer(infos, dtype=batch.INFO, count=m)["bytes"].astype(np.int64)
        padded = _up(ALIGN, np.maximum(sizes, 16))
        offs = np.zeros(m, np.uint64)
        offs[1:] = np.cumsum(padded)[:-1]
        host = torch.empty(int(padded.sum()), dtype=torch.uint8, pin_memory=True)
        check(lib.hgap_dectable_build_batch(arr, m, _vp(pb), infos,
                                            ctypes.c_void_p(host.data_ptr()), _vp(offs),
                                            host.numel(), 1), "hgap_dectable_build_batch")
        dev = host.to(self.device)
        for j, k in enumerate(ks):
            o, s = int(offs[j]), int(sizes[j])
            info = _capi.DectableInfo.from_buffer_copy(infos[j])
            t = self._put(k, codec.DeviceTable.resident(info, dev[o:o + s].view(torch.int32)))
            for i in miss[k]:
                out[i] = t
        return out


def _pinned_to_device(arr, device):
    host = torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8)).pin_memory()
    return host, host.to(device, non_blocking=True)


class _Batch:
    """One admitted batch in flight on one stream."""

    def __init__(self, server, reqs, stream):
        self.server = server
        self.reqs = reqs
        self.stream = stream
        self.enc = [r for r in reqs if r.kind == "encode"]
        self.dec = [r for r in reqs if r.kind == "decode"]
        self.keep = []  # host/device objects that must live until the batch completes
        self.hist_event = None
        self.done_event = None
        with torch.cuda.stream(stream):
            if self.dec:
                self._launch_decode()
            if self.enc:
                self._launch_histogram()
            else:
                self._finish_launches()

    # -- decode --
    def _launch_decode(self):
        s = self.server
        lib = s.lib
        groups = {}
        cfgs = s.dec_configs(self.dec)
        tables = s.tables.get_many([r.compressed.codebook for r in self.dec],
                                   [int(c.prefix_bits) for c in cfgs])
        for r, cfg, table in zip(self.dec, cfgs, tables):
            out = torch.empty(_up(16, r.n + 16), dtype=torch.uint8, device=s.device)
            r.result = out
            e = r.compressed
            expect = batch.expected_symbols(r.n, e.gap_elements,
                                            table.info.max_symbols_per_segment)
            key = (cfg.dec_threads, cfg.lookback_window, int(batch.density_class(expect)))
            groups.setdefault(key, (cfg, []))[1].append((r, table, out))
            self.keep.append(table)
        for cfg, members in groups.values():
            jobs = np.zeros(len(members), batch.DEC_JOB)
            max_tw, max_sym = 2, 1
            for j, (r, table, out) in enumerate(members):
                e, info = r.compressed, table.info
                jobs[j]["payload"] = e.payload.data_ptr()
                jobs[j]["gap"] = e.gap.data_ptr()
                jobs[j]["out"] = out.data_ptr()
                jobs[j]["n"] = r.n
                jobs[j]["table"] = table.table.data_ptr()
                jobs[j]["payload_words"] = e.payload_words
                jobs[j]["gap_elements"] = e.gap_elements
                jobs[j]["prefix_bits"] = info.prefix_bits
                jobs[j]["l1_entries"] = info.l1_entries
                jobs[j]["table_words"] = info.bytes // 4
                jobs[j]["max_symbols"] = info.max_symbols_per_segment
                jobs[j]["max_length"] = info.max_length
                max_tw = max(max_tw, info.bytes // 4)
                max_sym = max(max_sym, int(batch.expected_symbols(
                    r.n, e.gap_elements, info.max_symbols_per_segment)))
            total = lib.hgap_batch_plan_decode(_vp(jobs), len(members), ctypes.byref(cfg))
            host, d_jobs = _pinned_to_device(jobs, s.device)
            ws = torch.empty(_up(16, lib.hgap_batch_decode_workspace_bytes(total, ctypes.byref(cfg))),
                             dtype=torch.uint8, device=s.device)
            self.keep += [host, d_jobs, ws, cfg]
            check(lib.hgap_batch_decode(ctypes.c_void_p(d_jobs.data_ptr()), len(members), total,
                                        max_tw, max_sym, ctypes.c_void_p(ws.data_ptr()),
                                        ws.numel(), ctypes.byref(cfg), _sp(self.stream)),
                  "hgap_batch_decode")

    # -- encode, phase 1 --
    def _launch_histogram(self):
        s = self.server
        lib = s.lib
        self.hist = torch.empty((len(self.enc), 256), dtype=torch.int32, device=s.device)
        jobs = np.zeros(len(self.enc), batch.HIST_JOB)
        for j, r in enumerate(self.enc):
            jobs[j]["in_"] = r.data.data_ptr()
            jobs[j]["n"] = r.n
            jobs[j]["hist"] = self.hist.data_ptr() + 1024 * j
        lib.hgap_batch_plan_histogram(_vp(jobs), len(self.enc))
        # per-range counts too, for encode launches that take the chained encoder
        self.range_base = jobs["range_base"].astype(np.uint32).copy()
        ranges = int((-(-jobs["n"].astype(np.int64) // codec.RANGE_BYTES)).sum())
        self.range_hist = torch.empty((max(ranges, 1), 256), dtype=torch.int32, device=s.device)
        host, d_jobs = _pinned_to_device(jobs, s.device)
        self.keep += [host, d_jobs, self.range_hist]
        check(lib.hgap_batch_histogram_ranges(ctypes.c_void_p(d_jobs.data_ptr()), len(self.enc),
                                              ranges, ctypes.c_void_p(self.range_hist.data_ptr()),
                                              _sp(self.stream)), "hgap_batch_histogram_ranges")
        self.hist_host = torch.empty((len(self.enc), 256), dtype=torch.int32, pin_memory=True)
        self.hist_host.copy_(self.hist, non_blocking=True)
        self.hist_event = torch.cuda.Event()
        self.hist_event.record(self.stream)

    # -- encode, phase 2 (after the histogram completed) --
    def _launch_encode(self):
        s = self.server
        lib = s.lib
        hists = self.hist_host.numpy().view(np.uint32)
        cbs = batch.codebooks(hists, s.codebook_threads)
        groups = {}
        cfgs = s.enc_configs(hists, cbs)
        with torch.cuda.stream(self.stream):
            tables = np.zeros((len(self.enc), 256), np.uint32)
            for j, r in enumerate(self.enc):
                cb = cbs[j]
                sz = Sizes()
                check(lib.hgap_encode_sizes(ctypes.byref(cb), _vp(hists[j]), ctypes.byref(sz)),
                      "hgap_encode_sizes")
                exact = encode_reserve_exact(r.n, sz)
                s.resize(r, exact)
                # payload and gap array in one allocation (the gap starts 16-byte aligned)
                pw = _up(4, max(int(sz.payload_words), 4))
                words = torch.empty(pw + max(int(sz.gap_words), 4), dtype=torch.int32,
                                    device=s.device)
                r.result = codec.Encoded(r.n, cb, sz, words[:pw], words[pw:])
                check(lib.hgap_encode_table(ctypes.byref(cb), _vp(tables[j])),
                      "hgap_encode_table")
                cfg = cfgs[j]
                key = (cfg.enc_threads, cfg.enc_bytes_per_thread, cfg.lookback_window)
                groups.setdefault(key, (cfg, []))[1].append((j, r))
            host_t, d_tables = _pinned_to_device(tables, s.device)
            self.keep += [host_t, d_tables]
            for cfg, members in groups.values():
                jobs = np.zeros(len(members), batch.ENC_JOB)
                max_len = 1
                for q, (j, r) in enumerate(members):
                    e = r.result
                    jobs[q]["in_"] = r.data.data_ptr()
                    jobs[q]["n"] = r.n
                    jobs[q]["payload"] = e.payload.data_ptr()
                    jobs[q]["gap"] = e.gap.data_ptr()
                    jobs[q]["payload_words"] = e.payload_words
                    jobs[q]["gap_words"] = int(e.sizes.gap_words)
                    jobs[q]["table"] = j
                    ml = max(e.codebook.length[:]) or 1
                    jobs[q]["max_len"] = ml
                    jobs[q]["range_base"] = self.range_base[j]
                    max_len = max(max_len, ml)
                ranges = int((-(-jobs["n"].astype(np.int64) // codec.RANGE_BYTES)).sum())
                if ranges >= s.chained_min_ranges:
                    if s.default_enc_shape:
                        cfg = _capi.default_chained_config()
                    total = lib.hgap_batch_plan_encode_ranges(_vp(jobs), len(members),
                                                              ctypes.byref(cfg))
                    wsb = lib.hgap_batch_encode_ranges_workspace_bytes(total)
                else:
                    total = lib.hgap_batch_plan_encode(_vp(jobs), len(members), ctypes.byref(cfg))
                    wsb = lib.hgap_batch_workspace_bytes(total)
                host, d_jobs = _pinned_to_device(jobs, s.device)
                ws = torch.empty(_up(16, wsb), dtype=torch.uint8, device=s.device)
                self.keep += [host, d_jobs, ws, cfg]
                if ranges >= s.chained_min_ranges:
                    check(lib.hgap_batch_encode_ranges(
                        ctypes.c_void_p(d_jobs.data_ptr()), len(members), total, max_len,
                        ctypes.c_void_p(d_tables.data_ptr()),
                        ctypes.c_void_p(self.range_hist.data_ptr()),
                        ctypes.c_void_p(ws.data_ptr()), ws.numel(), ctypes.byref(cfg),
                        _sp(self.stream)), "hgap_batch_encode_ranges")
                else:
                    check(lib.hgap_batch_encode(ctypes.c_void_p(d_jobs.data_ptr()), len(members),
                                                total, max_len,
                                                ctypes.c_void_p(d_tables.data_ptr()),
                                                ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                                                ctypes.byref(cfg), _sp(self.stream)),
                          "hgap_batch_encode")
            self._finish_launches()

    def _finish_launches(self):
        self.done_event = torch.cuda.Event()
        self.done_event.record(self.stream)

    def poll(self):
        """Advances the batch; returns True once every request of it has completed."""
        if self.done_event is None:
            if not self.hist_event.query():
                return False
            self._launch_encode()
        return self.done_event.query()


class Server:
    """Continuous-batching request loop on one device."""

    def __init__(self, cap_bytes=200 * 10**9, max_batch=64, max_inflight=4, streams=2,
                 device="cuda", predictor=None, codebook_threads=0, config=None,
                 chained_min_ranges=batch.BatchCodec.CHAINED_MIN_RANGES):
        """An encode launch whose requests hold at least `chained_min_ranges` 256 KiB ranges
        takes the chained encoder (offsets from the range counts of the batched histogram)."""
        self.lib = _capi.load()
        self.chained_min_ranges = int(chained_min_ranges)
        self.device = device
        self.cap = int(cap_bytes)
        self.max_batch = int(max_batch)
        self.max_inflight = int(max_inflight)
        # default-priority streams only
        self.streams = [torch.cuda.Stream(device=device) for _ in range(max(1, streams))]
        self.tables = TableCache(device)
        self.predictor = predictor
        self.codebook_threads = codebook_threads
        self.config = LaunchConfig.from_dict(config or {})
        # without shapes from the caller, chained encode launches take the chained default
        self.default_enc_shape = config is None and predictor is None
        self.reserved = 0
        self.peak_reserved = 0
        self.batches = 0
        self.peak_inflight = 0
        self._next_stream = 0

    # launch shapes per request: the predictor's choice when one is given, else `config`
    def enc_configs(self, hists, cbs):
        if self.predictor is None:
            return [self.config] * len(cbs)
        from .predictor import choose_configs

        enc, _, _ = choose_configs(hists, cbs, self.predictor, fallback=False)
        return [LaunchConfig.from_dict(c) for c in enc]

    def dec_configs(self, reqs):
        if self.predictor is None:
            return [self.config] * len(reqs)
        from .predictor import choose_configs

        cbs = [r.compressed.codebook for r in reqs]
        _, dec, _ = choose_configs([implied_counts(r.compressed) for r in reqs], cbs,
                                   self.predictor, fallback=False)
        return [LaunchConfig.from_dict(c) for c in dec]

    def resize(self, r, nbytes):
        self.reserved += nbytes - r.reserved
        r.reserved = nbytes

    def _release(self, r):
        self.reserved -= r.reserved
        r.reserved = 0

    def run(self, requests, clock=time.perf_counter):
        """Serves `requests` (arrival times in seconds from the start); returns stats."""
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        queue = deque()
        inflight = []
        torch.cuda.synchronize(self.device)
        t0 = clock()
        while pending or queue or inflight:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            still = []
            for b in inflight:
                if b.poll():
                    t = clock() - t0
                    for r in b.reqs:
                        r.done = t
                        self._release(r)
                else:
                    still.append(b)
            inflight = still
            if queue and len(inflight) < self.max_inflight:
                reqs = []
                while queue and len(reqs) < self.max_batch:
                    r = queue[0]
                    need = request_bytes(r)
                    if need > self.cap:
                        queue.popleft()
                        r.error = f"request needs {need} bytes of device memory; cap {self.cap}"
                        r.done = clock() - t0
                        continue
                    if self.reserved + need > self.cap:
                        break
                    queue.popleft()
                    self.resize(r, need)
                    r.admitted = clock() - t0
                    reqs.append(r)
                self.peak_reserved = max(self.peak_reserved, self.reserved)
                if reqs:
                    for r in reqs:
                        r.batch_size = len(reqs)
                    stream = self.streams[self._next_stream % len(self.streams)]
                    self._next_stream += 1
                    inflight.append(_Batch(self, reqs, stream))
                    self.batches += 1
                    self.peak_inflight = max(self.peak_inflight, len(inflight))
                    continue
            if not inflight and not queue and pending:
                # idle until the next arrival
                wait = pending[0].arrival - (clock() - t0)
                if wait > 2e-4:
                    time.sleep(wait - 1e-4)
        torch.cuda.synchronize(self.device)
        return stats(requests)


def implied_counts(e):
    """Counts a decode request's features are computed from: a stream carries no histogram, so
    each symbol gets the count its code length implies for an optimal code, n * 2^-length."""
    h = np.zeros(256, np.float64)
    k = int(e.codebook.symbol_count)
    for j in range(k):
        h[e.codebook.order_symbol[j]] = max(1.0, e.n * 2.0 ** -int(e.codebook.order_length[j]))
    return h


def percentile(xs, q):
    if not xs:
        return None
    s = sorted(xs)
    i = min(len(s) - 1, max(0, int(math.ceil(q / 100.0 * len(s))) - 1))
    return s[i]


def stats(requests):
    served = [r for r in requests if r.error is None and r.done is not None]
    lat = [r.latency * 1e3 for r in served]
    out = dict(requests=len(requests), served=len(served),
               rejected=sum(1 for r in requests if r.error is not None))
    if not served:
        return out
    first = min(r.arrival for r in served)
    last = max(r.done for r in served)
    total = sum(r.n for r in served)
    out.update(makespan_s=round(last - first, 5), bytes=total,
               throughput_GBps=round(total / max(last - first, 1e-9) / GB, 2),
               requests_per_s=round(len(served) / max(last - first, 1e-9), 1),
               p50_ms=round(percentile(lat, 50), 4), p99_ms=round(percentile(lat, 99), 4),
               mean_ms=round(statistics.mean(lat), 4), max_ms=round(max(lat), 4),
               mean_batch=round(statistics.mean(r.batch_size for r in served), 2))
    for kind in ("encode", "decode"):
        lk = [r.latency * 1e3 for r in served if r.kind == kind]
        if lk:
            out[kind] = dict(count=len(lk), p50_ms=round(percentile(lk, 50), 4),
                             p99_ms=round(percentile(lk, 99), 4))
    return out


# ---- workload ---------------------------------------------------------------------------------

def workload_spec(count, seed=5, min_bytes=1 << 10, max_bytes=1 << 28):
    """Requests of the serving workload: kind 50/50, sizes log-uniform in [min, max] bytes,
    Zipf(s ~ U(0.6, 2.0)) data with its own permutation.  Deterministic in `seed`."""
    from . import synth

    out = []
    lo, hi = math.log(min_bytes), math.log(max_bytes)
    for i in range(count):
        u = [(synth.splitmix64_scalar(seed, 5 * i + j) >> 11) / float(synth.ONE53)
             for j in range(3)]
        out.append(dict(rid=i, kind="encode" if u[0] < 0.5 else "decode",
                        n=int(math.exp(lo + (hi - lo) * u[1])), s=0.6 + 1.4 * u[2],
                        seed=(seed << 32) + i, perm_seed=synth.splitmix64_scalar(seed, 5 * i + 4)))
    return out


def poisson_arrivals(count, rate, seed=17):
    """Arrival times (s) of a Poisson process of `rate` requests/s (rate = inf: all at 0)."""
    from . import synth

    if not math.isfinite(rate):
        return [0.0] * count
    t, out = 0.0, []
    for i in range(count):
        u = (synth.splitmix64_scalar(seed, i) >> 11) / float(synth.ONE53)
        t += -math.log1p(-u) / rate
        out.append(t)
    return out


def make_requests(spec, device="cuda"):
    """Device-resident requests: raw bytes for encode requests; for decode requests, the
    compressed words of the same kind of data (encoded here, before serving starts)."""
    from . import synth

    reqs = []
    ws = codec.Workspace(device)
    for d in spec:
        buf = torch.empty(_up(16, max(d["n"], 16)), dtype=torch.uint8, device=device)[:d["n"]]
        codec.synth_fill(buf, d["seed"], synth.cdf_from_pmf(synth.pmf_zipf(d["s"])),
                         synth.permutation(d["perm_seed"]))
        r = Request(d["rid"], d["kind"], d["n"], data=buf)
        if d["kind"] == "decode":
            r.compressed = codec.encode(buf, workspace=ws)
        reqs.append(r)
    torch.cuda.synchronize()
    return reqs


def reset(requests, arrivals=None):
    for i, r in enumerate(requests):
        r.result = r.admitted = r.done = r.error = None
        r.reserved = r.batch_size = 0
        if arrivals is not None:
            r.arrival = arrivals[i]


def verify(requests):
    """Mismatching requests: decode outputs against the original bytes, encode outputs by
    decoding them with a plain decode call."""
    bad = []
    ws = codec.Workspace(requests[0].data.device if requests else "cuda")
    for r in requests:
        if r.error is not None:
            continue
        if r.kind == "decode":
            ok = torch.equal(r.result[:r.n], r.data)
        else:
            e = r.result
            t = codec.DeviceTable(e.codebook, device=r.data.device)
            out = codec.decode(e.payload, e.gap, e.payload_words, e.gap_elements, t, r.n,
                               workspace=ws)
            ok = torch.equal(out[:r.n], r.data)
        if not ok:
            bad.append(r.rid)
    return bad


def bench(count=2000, load=0.65, seed=5, cap_bytes=200 * 10**9, max_batch=64, max_inflight=4,
          streams=2, device="cuda", log=None, rate=None):
    """Capacity (all requests queued at t = 0), then Poisson arrivals at `load` x capacity (or
    at `rate` requests per second when given)."""
    spec = workload_spec(count, seed)
    reqs = make_requests(spec, device)
    srv = lambda: Server(cap_bytes, max_batch, max_inflight, streams, device)  # noqa: E731
    reset(reqs, poisson_arrivals(count, math.inf))
    srv().run(reqs[: min(count, 64)])  # warm-up
    reset(reqs, poisson_arrivals(count, math.inf))
    sat = srv().run(reqs)
    bad_sat = verify(reqs)
    if rate is None:
        rate = load * sat["requests_per_s"]
    if log:
        log(f"capacity {sat['requests_per_s']} req/s, {sat['throughput_GBps']} GB/s; "
            f"offering {rate:.1f} req/s")
    reset(reqs, poisson_arrivals(count, rate))
    s = srv()
    res = s.run(reqs)
    bad = verify(reqs)
    res.update(offered_rate_per_s=round(rate, 1), load=load, capacity=sat,
               cap_bytes=cap_bytes, peak_reserved_bytes=s.peak_reserved, batches=s.batches,
               peak_inflight_batches=s.peak_inflight,
               table_cache=dict(hits=s.tables.hits, misses=s.tables.misses),
               mismatches=len(bad) + len(bad_sat),
               workload=(f"{count} requests, 50/50 encode/decode, sizes log-uniform "
                         "1 KiB..256 MiB, Zipf s ~ U(0.6, 2.0), Poisson arrivals"))
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m huffman_coding_gap_arrays_amd.serve")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--load", type=float, default=0.65)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--cap-gb", type=float, default=200.0)
    ap.add_argument("--max-batch", type=int, default=64)
    ap.add_argument("--inflight", type=int, default=4)
    ap.add_argument("--streams", type=int, default=2)
    ap.add_argument("--rate", type=float, default=None,
                    help="offered requests per second (default: --load x measured capacity)")
    args = ap.parse_args(argv)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)  # noqa: E731
    res = bench(args.requests, args.load, args.seed, int(args.cap_gb * 1e9), args.max_batch,
                args.inflight, args.streams, log=log, rate=args.rate)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

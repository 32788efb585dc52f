"""Batched encode/decode of many independent buffers, each with its own codebook.

All buffers of a batch are processed by a few persistent launches (one batched histogram, then
one encode and one decode launch per launch-shape group), driven by job descriptors resident in
device memory.  Because nothing on the device path depends on host values once the descriptors
are uploaded, the encode + decode launches can be captured once in a HIP graph
(torch.cuda.CUDAGraph) and replayed.

Per-buffer launch shapes (encoder tile shape, decoder tile size, decode-table split) come from a
`configs` list -- the exhaustive autotuner's or the learned predictor's choices -- and buffers
sharing a shape share a launch.
"""
import ctypes
from collections import defaultdict

import numpy as np
import torch

from . import _capi, codec
from ._capi import Codebook, DecJob, DectableInfo, EncJob, HistJob, LaunchConfig, Sizes, check

HIST_JOB = np.dtype(HistJob)
ENC_JOB = np.dtype(EncJob)
DEC_JOB = np.dtype(DecJob)
ALIGN = 256  # arena alignment of every buffer (bytes)


def _lib():
    return _capi.load()


def _vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _up(b, x):
    return (x + b - 1) // b * b


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _to_device(arr, device):
    return torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).copy()).to(device)


class Arena:
    """One device allocation carved into 256-byte aligned views."""

    def __init__(self, sizes, device="cuda"):
        self.sizes = [int(s) for s in sizes]
        self.offsets = []
        off = 0
        for s in self.sizes:
            self.offsets.append(off)
            off = _up(ALIGN, off + max(s, 1))
        self.mem = torch.empty(max(off, ALIGN), dtype=torch.uint8, device=device)
        self.base = self.mem.data_ptr()

    def view(self, i):
        o, s = self.offsets[i], self.sizes[i]
        return self.mem[o:o + s]

    def ptr(self, i):
        return self.base + self.offsets[i]


def histograms(buffers, stream=None, device="cuda"):
    """256 counts per buffer (uint32 [B, 256] on the host), one batched launch."""
    B = len(buffers)
    hist = torch.empty((max(B, 1), 256), dtype=torch.int32, device=device)
    jobs = np.zeros(B, HIST_JOB)
    for i, b in enumerate(buffers):
        jobs[i]["in_"] = b.data_ptr()
        jobs[i]["n"] = b.numel()
        jobs[i]["hist"] = hist.data_ptr() + 1024 * i
    total = _lib().hgap_batch_plan_histogram(_vp(jobs), B)
    d_jobs = _to_device(jobs, device)
    check(_lib().hgap_batch_histogram(ctypes.c_void_p(d_jobs.data_ptr()), B, total,
                                      _stream(stream)), "hgap_batch_histogram")
    return hist[:B].cpu().numpy().view(np.uint32)


def codebooks(hists, threads=0):
    """Codebooks of [B, 256] histograms, built on a pool of host threads."""
    h = np.ascontiguousarray(hists, dtype=np.uint32)
    cbs = (Codebook * max(len(h), 1))()
    check(_lib().hgap_codebook_build_batch(_vp(h), len(h), cbs, threads),
          "hgap_codebook_build_batch")
    return [cbs[i] for i in range(len(h))]


def _cfg_struct(cfg):
    return LaunchConfig.from_dict(cfg or {})


class BatchCodec:
    """Encodes and decodes a list of device uint8 buffers in batched launches."""

    def __init__(self, buffers, configs=None, hists=None, device="cuda", stream=None,
                 threads=0, dec_configs=None, predictor=None):
        """configs[i] / dec_configs[i]: launch shapes of buffer i's encode / decode (dicts of
        hgap_launch_config fields; dec_configs defaults to configs).  With a `predictor`
        (predictor.Predictor), both come from its per-buffer choice instead."""
        lib = _lib()
        self.device = device
        self.buffers = list(buffers)
        B = len(self.buffers)
        if hists is None:
            hists = histograms(self.buffers, stream, device)
        self.hists = np.ascontiguousarray(hists, dtype=np.uint32)
        self.codebooks = codebooks(self.hists, threads)
        if predictor is not None:
            from .predictor import choose_configs

            configs, dec_configs, _ = choose_configs(self.hists, self.codebooks, predictor,
                                                     fallback=False)
        self.configs = [dict(c) for c in (configs or [{}] * B)]
        self.dec_configs = [dict(c) for c in (dec_configs or self.configs)]
        self.sizes = []
        for i in range(B):
            sz = Sizes()
            check(lib.hgap_encode_sizes(ctypes.byref(self.codebooks[i]), _vp(self.hists[i]),
                                        ctypes.byref(sz)), "hgap_encode_sizes")
            self.sizes.append(sz)
        ns = [b.numel() for b in self.buffers]
        self.payload = Arena([4 * s.payload_words for s in self.sizes], device)
        self.gap = Arena([4 * s.gap_words for s in self.sizes], device)
        self.out = Arena(ns, device)
        # encoder code tables, one per buffer
        tables = np.zeros((max(B, 1), 256), np.uint32)
        for i, cb in enumerate(self.codebooks):
            check(lib.hgap_encode_table(ctypes.byref(cb), _vp(tables[i])), "hgap_encode_table")
        self.d_enc_tables = _to_device(tables, device)
        # decode tables, one per buffer (first-level width from its config)
        infos, images = [], []
        for i, cb in enumerate(self.codebooks):
            info = DectableInfo()
            p = int(self.dec_configs[i].get("prefix_bits", 0))
            check(lib.hgap_dectable_build(ctypes.byref(cb), p, ctypes.byref(info), None, 0),
                  "hgap_dectable_build")
            img = np.zeros(max(info.bytes // 4, 4), np.uint32)
            if info.bytes:
                check(lib.hgap_dectable_build(ctypes.byref(cb), p, ctypes.byref(info), _vp(img),
                                              img.nbytes), "hgap_dectable_build")
            infos.append(info)
            images.append(img)
        self.dec_infos = infos
        self.dec_tables = Arena([im.nbytes for im in images], device)
        host = np.zeros(self.dec_tables.mem.numel(), np.uint8)
        for i, im in enumerate(images):
            o = self.dec_tables.offsets[i]
            host[o:o + im.nbytes] = im.view(np.uint8)
        self.dec_tables.mem.copy_(torch.from_numpy(host))
        self._plan_groups()

    def _plan_groups(self):
        lib = _lib()
        enc_groups = defaultdict(list)
        dec_groups = defaultdict(list)
        for i, b in enumerate(self.buffers):
            if b.numel() == 0:
                continue
            c = _cfg_struct(self.configs[i])
            enc_groups[(c.enc_threads, c.enc_bytes_per_thread, c.lookback_window)].append(i)
            d = _cfg_struct(self.dec_configs[i])
            # one LDS slot size per launch: every member's shape then fits when it fits alone
            slot = lib.hgap_decode_slot_words(self.dec_infos[i].max_symbols_per_segment)
            dec_groups[(d.dec_threads, d.lookback_window, slot)].append(i)
        self.enc_launches = []
        for (t, k, w), idx in sorted(enc_groups.items()):
            cfg = _cfg_struct(dict(enc_threads=t, enc_bytes_per_thread=k, lookback_window=w))
            jobs = np.zeros(len(idx), ENC_JOB)
            max_len = 1
            for j, i in enumerate(idx):
                sz = self.sizes[i]
                jobs[j]["in_"] = self.buffers[i].data_ptr()
                jobs[j]["n"] = self.buffers[i].numel()
                jobs[j]["payload"] = self.payload.ptr(i)
                jobs[j]["gap"] = self.gap.ptr(i)
                jobs[j]["payload_words"] = sz.payload_words
                jobs[j]["gap_words"] = sz.gap_words
                jobs[j]["table"] = i
                ml = max(self.codebooks[i].length[:]) or 1
                jobs[j]["max_len"] = ml
                max_len = max(max_len, ml)
            total = lib.hgap_batch_plan_encode(_vp(jobs), len(idx), ctypes.byref(cfg))
            ws = torch.empty(_up(16, lib.hgap_batch_workspace_bytes(total)), dtype=torch.uint8,
                             device=self.device)
            self.enc_launches.append(dict(cfg=cfg, jobs=_to_device(jobs, self.device),
                                          count=len(idx), total=total, max_len=max_len, ws=ws,
                                          members=idx))
        self.dec_launches = []
        for (t, w, _), idx in sorted(dec_groups.items()):
            cfg = _cfg_struct(dict(dec_threads=t, lookback_window=w))
            jobs = np.zeros(len(idx), DEC_JOB)
            max_tw, max_sym = 2, 1
            for j, i in enumerate(idx):
                sz, info = self.sizes[i], self.dec_infos[i]
                jobs[j]["payload"] = self.payload.ptr(i)
                jobs[j]["gap"] = self.gap.ptr(i)
                jobs[j]["out"] = self.out.ptr(i)
                jobs[j]["n"] = self.buffers[i].numel()
                jobs[j]["table"] = self.dec_tables.ptr(i)
                jobs[j]["payload_words"] = sz.payload_words
                jobs[j]["gap_elements"] = sz.gap_elements
                jobs[j]["prefix_bits"] = info.prefix_bits
                jobs[j]["l1_entries"] = info.l1_entries
                jobs[j]["table_words"] = info.l1_entries + info.l2_entries
                jobs[j]["max_symbols"] = info.max_symbols_per_segment
                jobs[j]["max_length"] = info.max_length
                max_tw = max(max_tw, info.l1_entries + info.l2_entries)
                max_sym = max(max_sym, info.max_symbols_per_segment)
            total = lib.hgap_batch_plan_decode(_vp(jobs), len(idx), ctypes.byref(cfg))
            ws = torch.empty(_up(16, lib.hgap_batch_workspace_bytes(total)), dtype=torch.uint8,
                             device=self.device)
            self.dec_launches.append(dict(cfg=cfg, jobs=_to_device(jobs, self.device),
                                          count=len(idx), total=total, max_tw=max_tw,
                                          max_sym=max_sym, ws=ws, members=idx))

    def encode(self, stream=None):
        lib = _lib()
        for L in self.enc_launches:
            check(lib.hgap_batch_encode(ctypes.c_void_p(L["jobs"].data_ptr()), L["count"],
                                        L["total"], L["max_len"],
                                        ctypes.c_void_p(self.d_enc_tables.data_ptr()),
                                        ctypes.c_void_p(L["ws"].data_ptr()), L["ws"].numel(),
                                        ctypes.byref(L["cfg"]), _stream(stream)),
                  "hgap_batch_encode")

    def decode(self, stream=None):
        lib = _lib()
        for L in self.dec_launches:
            check(lib.hgap_batch_decode(ctypes.c_void_p(L["jobs"].data_ptr()), L["count"],
                                        L["total"], L["max_tw"], L["max_sym"],
                                        ctypes.c_void_p(L["ws"].data_ptr()), L["ws"].numel(),
                                        ctypes.byref(L["cfg"]), _stream(stream)),
                  "hgap_batch_decode")

    def run(self, stream=None):
        self.encode(stream)
        self.decode(stream)

    def capture(self):
        """A HIP graph of encode + decode for this batch (replay with .replay())."""
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.run(side)  # warm up: kernel attributes and occupancy queries outside capture
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            self.run(torch.cuda.current_stream())
        return g

    def payload_u32(self, i):
        return self.payload.view(i).view(torch.int32).cpu().numpy().view(np.uint32)

    def gap_u32(self, i):
        return self.gap.view(i).view(torch.int32).cpu().numpy().view(np.uint32)

    def output(self, i):
        return self.out.view(i)

    def mismatches(self):
        """Number of buffers whose decoded bytes differ from the input."""
        return sum(0 if torch.equal(self.out.view(i), b) else 1
                   for i, b in enumerate(self.buffers))

    @property
    def total_bytes(self):
        return sum(b.numel() for b in self.buffers)

    @property
    def compressed_bytes(self):
        return sum(4 * (s.payload_words + s.gap_words) for s in self.sizes)


def plain_loop(buffers, stream=None, workspace=None):
    """The per-buffer path: histogram, codebook, encode, table, decode, one buffer at a time
    (what a caller of the plain entry points does); returns the decoded buffers."""
    ws = workspace or codec.Workspace(buffers[0].device if buffers else "cuda")
    outs = []
    for b in buffers:
        if b.numel() == 0:
            outs.append(b)
            continue
        enc = codec.encode(b, stream=stream, workspace=ws)
        table = codec.DeviceTable(enc.codebook, device=b.device)
        outs.append(codec.decode(enc.payload, enc.gap, enc.payload_words, enc.gap_elements, table,
                                 enc.n, stream=stream, workspace=ws))
    return outs


def make_workload(count=4096, seed=4, device="cuda", **kw):
    """(arena, buffers, spec) of the batched workload, generated on the device."""
    from . import synth

    spec = synth.batch_spec(count, seed, **kw)
    arena = Arena([d["n"] for d in spec], device)
    bufs = []
    for i, d in enumerate(spec):
        b = arena.view(i)
        codec.synth_fill(b, d["seed"], synth.cdf_from_pmf(synth.pmf_zipf(d["s"])),
                         synth.permutation(d["perm_seed"]))
        bufs.append(b)
    torch.cuda.synchronize()
    return arena, bufs, spec


def bench(count=4096, seed=4, reps=3, device="cuda", variants=(("default", None, None),)):
    """Wall-clock comparison on the batched workload (all inputs resident in HBM):
    plain   - per-buffer histogram/codebook/encode/table/decode calls, one buffer at a time;
    batched - batched histogram, threaded codebooks, launch-shape choice, planning and table
              upload, batched encode and decode (the whole pipeline from raw buffers);
    graph   - replay of the captured encode + decode graph (codebooks and tables reused; the
              payload and output arenas are cleared first, so the check after the replays sees
              only what they wrote).
    `variants`: (name, configs, predictor) per batched run; the launch shapes come from
    `configs` (None = defaults) or from the predictor's per-buffer choice."""
    import time

    arena, bufs, spec = make_workload(count, seed, device)
    total = sum(b.numel() for b in bufs)

    def timed(fn):
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    ws = codec.Workspace(device)
    plain_outs = plain_loop(bufs, workspace=ws)  # warm-up + correctness
    plain_bad = sum(0 if torch.equal(o, b) else 1 for o, b in zip(plain_outs, bufs))
    del plain_outs
    t_plain = timed(lambda: plain_loop(bufs, workspace=ws))
    out = dict(buffers=count, total_bytes=total, plain_s=round(t_plain, 5),
               plain_GBps=round(total / t_plain / 1e9, 2), plain_mismatches=int(plain_bad))
    for name, configs, predictor in variants:
        holder = {}

        def batched():
            bc = BatchCodec(bufs, configs=configs, device=device, predictor=predictor)
            bc.run()
            holder["bc"] = bc

        t_batched = timed(batched)
        bc = holder["bc"]
        bad = bc.mismatches()
        g = bc.capture()
        bc.payload.mem.zero_()
        bc.out.mem.zero_()
        t_graph = timed(g.replay)
        bad_graph = bc.mismatches()
        shapes = len({(c.get("enc_threads"), c.get("enc_bytes_per_thread"))
                      for c in bc.configs})
        out[name] = dict(batched_s=round(t_batched, 5), graph_s=round(t_graph, 5),
                         batched_GBps=round(total / t_batched / 1e9, 2),
                         graph_GBps=round(total / t_graph / 1e9, 2),
                         speedup_batched_vs_plain=round(t_plain / t_batched, 2),
                         speedup_graph_vs_plain=round(t_plain / t_graph, 2),
                         encode_launches=len(bc.enc_launches),
                         decode_launches=len(bc.dec_launches), encoder_tile_shapes=shapes,
                         compressed_bytes=bc.compressed_bytes, mismatches=int(bad + bad_graph))
        del g, bc, holder
    return out

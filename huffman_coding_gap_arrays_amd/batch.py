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

import numpy as np
import torch

from . import _capi, codec
from ._capi import Codebook, DecJob, DectableInfo, EncJob, HistJob, LaunchConfig, Sizes, check

HIST_JOB = np.dtype(HistJob)
ENC_JOB = np.dtype(EncJob)
DEC_JOB = np.dtype(DecJob)
SIZES = np.dtype(Sizes)
INFO = np.dtype(DectableInfo)
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
        self.sizes = np.asarray(sizes, dtype=np.int64).reshape(-1)
        padded = _up(ALIGN, np.maximum(self.sizes, 1))
        self.offsets = np.zeros(len(self.sizes), np.int64)
        if len(padded) > 1:
            self.offsets[1:] = np.cumsum(padded)[:-1]
        total = int(padded.sum()) if len(padded) else 0
        self.mem = torch.empty(max(total, ALIGN), dtype=torch.uint8, device=device)
        self.base = self.mem.data_ptr()

    def view(self, i):
        o, s = int(self.offsets[i]), int(self.sizes[i])
        return self.mem[o:o + s]

    def ptr(self, i):
        return self.base + int(self.offsets[i])

    def ptrs(self, idx=slice(None)):
        return (self.offsets[idx] + self.base).astype(np.uint64)


def histograms(buffers, stream=None, device="cuda"):
    """256 counts per buffer (uint32 [B, 256] on the host), one batched launch."""
    B = len(buffers)
    hist = torch.empty((max(B, 1), 256), dtype=torch.int32, device=device)
    jobs = np.zeros(B, HIST_JOB)
    jobs["in_"] = [b.data_ptr() for b in buffers]
    jobs["n"] = [b.numel() for b in buffers]
    jobs["hist"] = hist.data_ptr() + 1024 * np.arange(B, dtype=np.uint64)
    total = _lib().hgap_batch_plan_histogram(_vp(jobs), B)
    d_jobs = _to_device(jobs, device)
    check(_lib().hgap_batch_histogram(ctypes.c_void_p(d_jobs.data_ptr()), B, total,
                                      _stream(stream)), "hgap_batch_histogram")
    return hist[:B].cpu().numpy().view(np.uint32)


def histograms_ranges(buffers, stream=None, device="cuda"):
    """(hists, range_hist, range_base): the 256 counts per buffer (uint32 [B, 256] on the host)
    and the counts of every 256 KiB range (int32 device tensor, one row per range; buffer i's
    rows start at range_base[i]), in one batched launch."""
    B = len(buffers)
    hist = torch.empty((max(B, 1), 256), dtype=torch.int32, device=device)
    jobs = np.zeros(B, HIST_JOB)
    jobs["in_"] = [b.data_ptr() for b in buffers]
    jobs["n"] = [b.numel() for b in buffers]
    jobs["hist"] = hist.data_ptr() + 1024 * np.arange(B, dtype=np.uint64)
    _lib().hgap_batch_plan_histogram(_vp(jobs), B)
    ranges = -(-jobs["n"].astype(np.int64) // codec.RANGE_BYTES)
    total = int(ranges.sum())
    range_hist = torch.empty((max(total, 1), 256), dtype=torch.int32, device=device)
    d_jobs = _to_device(jobs, device)
    check(_lib().hgap_batch_histogram_ranges(ctypes.c_void_p(d_jobs.data_ptr()), B, total,
                                             ctypes.c_void_p(range_hist.data_ptr()),
                                             _stream(stream)), "hgap_batch_histogram_ranges")
    return (hist[:B].cpu().numpy().view(np.uint32), range_hist,
            jobs["range_base"].astype(np.uint32).copy())


def codebooks(hists, threads=0):
    """Codebooks of [B, 256] histograms, built on a pool of host threads."""
    h = np.ascontiguousarray(hists, dtype=np.uint32)
    cbs = (Codebook * max(len(h), 1))()
    check(_lib().hgap_codebook_build_batch(_vp(h), len(h), cbs, threads),
          "hgap_codebook_build_batch")
    return [cbs[i] for i in range(len(h))]


def _cfg_struct(cfg):
    return LaunchConfig.from_dict(cfg or {})


def _config_arrays(configs, B):
    """Per-buffer launch-shape fields (dict of uint32 arrays) from a list of config dicts."""
    d = _capi.default_config().as_dict()
    out = {k: np.full(B, v, np.uint32) for k, v in d.items()}
    if configs:
        for i, c in enumerate(configs):
            for k, v in c.items():
                out[k][i] = v
    return out


def expected_symbols(n, gap_elements, max_symbols):
    """Codewords per segment the batched decoder sizes its LDS image for: 1.25 x the stream's
    mean + 2, at most the table's worst case (denser tiles take a slower path)."""
    n = np.asarray(n, np.int64)
    g = np.maximum(np.asarray(gap_elements, np.int64), 1)
    return np.minimum(np.asarray(max_symbols, np.int64), (5 * n + 4 * g - 1) // (4 * g) + 2)


# Decode launches are split at these expected codewords per segment: one launch sizes its LDS
# image for its densest member, so sparse buffers go in launches of their own and keep the
# occupancy their density allows (at most 32, 33-64 and more than 64).
DENSITY_CLASSES = (32, 64)


def density_class(expected):
    return np.searchsorted(DENSITY_CLASSES, np.asarray(expected), side="left")


class BatchCodec:
    """Encodes and decodes a list of device uint8 buffers in batched launches."""

    # an encode launch with at least this many 256 KiB ranges takes the chained encoder
    CHAINED_MIN_RANGES = 512

    def __init__(self, buffers, configs=None, hists=None, device="cuda", stream=None,
                 threads=0, dec_configs=None, predictor=None, chained=None):
        """configs[i] / dec_configs[i]: launch shapes of buffer i's encode / decode (dicts of
        hgap_launch_config fields; dec_configs defaults to configs).  With a `predictor`
        (predictor.Predictor), both come from its per-buffer choice instead.  Host preparation
        (codebooks, sizes, code tables, decode tables) runs on `threads` host threads.
        Without `hists`, the batched histogram also counts every 256 KiB range, and encode
        launches use the chained encoder: when they have CHAINED_MIN_RANGES ranges or more
        (chained=None), always (True) or never (False)."""
        from . import predictor as pr

        lib = _lib()
        self.device = device
        self.buffers = list(buffers)
        B = self.count = len(self.buffers)
        self.range_hist = self.range_base = None
        self.chained = chained
        if hists is None:
            if chained is False:
                hists = histograms(self.buffers, stream, device)
            else:
                hists, self.range_hist, self.range_base = histograms_ranges(self.buffers, stream,
                                                                            device)
        self.hists = np.ascontiguousarray(hists, dtype=np.uint32).reshape(B, 256)
        self.cb_array = (Codebook * max(B, 1))()
        self.size_array = (Sizes * max(B, 1))()
        tables = np.zeros((max(B, 1), 256), np.uint32)
        check(lib.hgap_batch_prepare(_vp(self.hists), B, self.cb_array, self.size_array,
                                     _vp(tables), threads), "hgap_batch_prepare")
        sv = np.frombuffer(self.size_array, dtype=SIZES, count=B)
        lengths = pr.code_lengths(self.cb_array, B)
        self.max_len = np.maximum(lengths.max(1) if B else np.zeros(0, np.uint32), 1)
        if predictor is not None and B:
            X = pr.features_batch(self.hists, lengths)
            em, dm = pr.support_masks(self.cb_array, B, self.max_len, threads)
            ei, di = pr.choose_indices(predictor, X, em, dm)
            configs = [pr.ENC_CANDIDATES[j] for j in ei]
            dec_configs = [pr.DEC_CANDIDATES[j] for j in di]
        # without shapes from the caller, chained encode launches take the chained default
        self.default_enc_shape = configs is None
        self.configs = [dict(c) for c in (configs or [{}] * B)]
        self.dec_configs = [dict(c) for c in (dec_configs or self.configs)]
        self.enc_cfg = _config_arrays(self.configs, B)
        self.dec_cfg = _config_arrays(self.dec_configs, B)
        self.ns = np.array([b.numel() for b in self.buffers], np.uint64)
        self.in_ptrs = np.array([b.data_ptr() for b in self.buffers], np.uint64)
        self.payload_words = sv["payload_words"].copy()
        self.gap_words = sv["gap_words"].copy()
        self.gap_elements = sv["gap_elements"].copy()
        self.payload = Arena(4 * self.payload_words.astype(np.int64), device)
        self.gap = Arena(4 * self.gap_words.astype(np.int64), device)
        self.out = Arena(self.ns.astype(np.int64), device)
        self.d_enc_tables = _to_device(tables, device)
        # decode tables, one per buffer, first-level width from its decoder config
        prefix = np.ascontiguousarray(self.dec_cfg["prefix_bits"])
        infos = (DectableInfo * max(B, 1))()
        check(lib.hgap_dectable_build_batch(self.cb_array, B, _vp(prefix), infos, None, None, 0,
                                            threads), "hgap_dectable_build_batch")
        iv = np.frombuffer(infos, dtype=INFO, count=B)
        self.dec_tables = Arena(iv["bytes"].astype(np.int64), device)
        host = np.zeros(self.dec_tables.mem.numel(), np.uint8)
        offs = np.ascontiguousarray(self.dec_tables.offsets.astype(np.uint64))
        check(lib.hgap_dectable_build_batch(self.cb_array, B, _vp(prefix), infos, _vp(host),
                                            _vp(offs), host.nbytes, threads),
              "hgap_dectable_build_batch")
        self.dec_info = np.frombuffer(infos, dtype=INFO, count=B).copy()
        self.dec_tables.mem.copy_(torch.from_numpy(host))
        self._plan_groups()

    @property
    def codebooks(self):
        return [self.cb_array[i] for i in range(self.count)]

    @property
    def sizes(self):
        return [self.size_array[i] for i in range(self.count)]

    def _workspace(self, total, dec_cfg=None):
        """Encode workspace for `total` tiles, or the decode workspace of shape `dec_cfg`."""
        lib = _lib()
        nbytes = (lib.hgap_batch_workspace_bytes(total) if dec_cfg is None
                  else lib.hgap_batch_decode_workspace_bytes(total, ctypes.byref(dec_cfg)))
        return torch.empty(_up(16, nbytes), dtype=torch.uint8, device=self.device)

    def _plan_groups(self):
        lib = _lib()
        live = np.nonzero(self.ns > 0)[0]
        e, d, iv = self.enc_cfg, self.dec_cfg, self.dec_info
        self.enc_launches = []
        ekeys = np.stack([e["enc_threads"], e["enc_bytes_per_thread"], e["lookback_window"]], 1)
        for key in np.unique(ekeys[live], axis=0):
            idx = live[(ekeys[live] == key).all(1)]
            cfg = _cfg_struct(dict(enc_threads=key[0], enc_bytes_per_thread=key[1],
                                   lookback_window=key[2]))
            jobs = np.zeros(len(idx), ENC_JOB)
            jobs["in_"] = self.in_ptrs[idx]
            jobs["n"] = self.ns[idx]
            jobs["payload"] = self.payload.ptrs(idx)
            jobs["gap"] = self.gap.ptrs(idx)
            jobs["payload_words"] = self.payload_words[idx]
            jobs["gap_words"] = self.gap_words[idx]
            jobs["table"] = idx
            jobs["max_len"] = self.max_len[idx]
            chained = False
            if self.range_hist is not None and self.chained is not False:
                ranges = int((-(-self.ns[idx].astype(np.int64) // codec.RANGE_BYTES)).sum())
                chained = self.chained or ranges >= self.CHAINED_MIN_RANGES
            if chained:
                if self.default_enc_shape:
                    dc = _capi.default_chained_config()
                    cfg.enc_threads = dc.enc_threads
                    cfg.enc_bytes_per_thread = dc.enc_bytes_per_thread
                jobs["range_base"] = self.range_base[idx]
                total = lib.hgap_batch_plan_encode_ranges(_vp(jobs), len(idx), ctypes.byref(cfg))
                ws = torch.empty(_up(16, lib.hgap_batch_encode_ranges_workspace_bytes(total)),
                                 dtype=torch.uint8, device=self.device)
            else:
                total = lib.hgap_batch_plan_encode(_vp(jobs), len(idx), ctypes.byref(cfg))
                ws = self._workspace(total)
            self.enc_launches.append(dict(cfg=cfg, jobs=_to_device(jobs, self.device),
                                          count=len(idx), total=total, chained=chained,
                                          max_len=int(self.max_len[idx].max()),
                                          ws=ws, members=idx))
        self.dec_launches = []
        ms = iv["max_symbols_per_segment"].astype(np.int64)
        tw = iv["bytes"].astype(np.int64) // 4  # words of each table image
        # one launch per decoder shape and density class; its LDS image is sized from the
        # members' largest table and densest segments (tiles that outgrow it go through the
        # region in windows)
        expect = expected_symbols(self.ns, self.gap_elements, ms)
        dkeys = np.stack([d["dec_threads"], d["lookback_window"], density_class(expect)], 1)
        for key in np.unique(dkeys[live], axis=0):
            idx = live[(dkeys[live] == key).all(1)]
            cfg = _cfg_struct(dict(dec_threads=key[0], lookback_window=key[1]))
            jobs = np.zeros(len(idx), DEC_JOB)
            jobs["payload"] = self.payload.ptrs(idx)
            jobs["gap"] = self.gap.ptrs(idx)
            jobs["out"] = self.out.ptrs(idx)
            jobs["n"] = self.ns[idx]
            jobs["table"] = self.dec_tables.ptrs(idx)
            jobs["payload_words"] = self.payload_words[idx]
            jobs["gap_elements"] = self.gap_elements[idx]
            jobs["prefix_bits"] = iv["prefix_bits"][idx]
            jobs["l1_entries"] = iv["l1_entries"][idx]
            jobs["table_words"] = tw[idx]
            jobs["max_symbols"] = ms[idx]
            dense = expect[idx]
            jobs["max_length"] = iv["max_length"][idx]
            total = lib.hgap_batch_plan_decode(_vp(jobs), len(idx), ctypes.byref(cfg))
            self.dec_launches.append(dict(cfg=cfg, jobs=_to_device(jobs, self.device),
                                          count=len(idx), total=total,
                                          max_tw=int(max(2, tw[idx].max())),
                                          max_sym=int(max(1, dense.max())),
                                          ws=self._workspace(total, cfg), members=idx))

    def encode(self, stream=None):
        lib = _lib()
        for L in self.enc_launches:
            if L["chained"]:
                check(lib.hgap_batch_encode_ranges(
                    ctypes.c_void_p(L["jobs"].data_ptr()), L["count"], L["total"], L["max_len"],
                    ctypes.c_void_p(self.d_enc_tables.data_ptr()),
                    ctypes.c_void_p(self.range_hist.data_ptr()),
                    ctypes.c_void_p(L["ws"].data_ptr()), L["ws"].numel(),
                    ctypes.byref(L["cfg"]), _stream(stream)), "hgap_batch_encode_ranges")
                continue
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

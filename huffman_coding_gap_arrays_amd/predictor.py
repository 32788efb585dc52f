"""Learned launch-shape predictor.

Tiny MLPs map five features of a buffer -- histogram entropy, size, longest and mean code
length, symbol count -- to the predicted relative run time of every candidate launch shape of
the encoder (27: workgroup size x bytes per thread x look-back window) and of the decoder (45:
workgroup size x look-back window x first-level table bits).  The prediction is the mean of an
ensemble of such networks trained from different seeds.  The shape with the lowest prediction
that the device can hold (hgap_config_supported) is used.

Training data are the exhaustive autotuner's sweeps (tuner.py): for every buffer of a sweep, the
median time of every candidate.  The regression target of candidate c is log2(t_c / t_best)
(0 for the best); a shape the device cannot run gets a fixed penalty.  The network is trained on
the GPU by libhgap's own kernels (hgap_mlp_forward / hgap_mlp_backward / hgap_mlp_adam_step);
there is no host path for it.  Exhaustive tuning is the fallback for buffers whose features
fall outside the training range, and the judge of the predictor's regret:
    regret(choice) = sum_i t_i(choice_i) / sum_i t_i(best_i) - 1      over held-out buffers.

usage:
  python -m huffman_coding_gap_arrays_amd.predictor sweep --count 512 --out sweeps.npz
  python -m huffman_coding_gap_arrays_amd.predictor train --sweeps sweeps.npz --out model.npz
  python -m huffman_coding_gap_arrays_amd.predictor eval  --sweeps sweeps.npz --model model.npz
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

from . import _capi
from ._capi import DectableInfo, LaunchConfig, check

FEATURE_NAMES = ("entropy_bits", "log2_bytes", "max_length", "mean_length", "log2_symbols")
UNSUPPORTED_PENALTY = 3.0  # log2 of the cost assigned to shapes the device cannot run (8x)
HIDDEN = 32
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_MODEL = os.path.join(DATA_DIR, "predictor.npz")
DEFAULT_SWEEPS = os.path.join(DATA_DIR, "launch_sweeps.npz")


def _space():
    from . import tuner

    return list(tuner.enc_space()), list(tuner.dec_space())


ENC_CANDIDATES, DEC_CANDIDATES = _space()
N_ENC, N_DEC = len(ENC_CANDIDATES), len(DEC_CANDIDATES)
N_OUT = N_ENC + N_DEC

DEFAULT_ENC = dict(enc_threads=512, enc_bytes_per_thread=64, lookback_window=64)
DEFAULT_DEC = dict(dec_threads=256, lookback_window=64, prefix_bits=11)


def candidate_index(cfg, candidates):
    keys = candidates[0].keys()
    for i, c in enumerate(candidates):
        if all(int(cfg[k]) == c[k] for k in keys):
            return i
    raise KeyError(cfg)


# ---- features --------------------------------------------------------------------------------

def features_batch(hists, lengths):
    """Feature rows [B, 5] (float64) from counts [B, 256] and code lengths by byte [B, 256]."""
    h = np.atleast_2d(np.asarray(hists, dtype=np.float64))
    ln = np.atleast_2d(np.asarray(lengths, dtype=np.float64))
    n = h.sum(1)
    safe_n = np.where(n > 0, n, 1.0)
    p = h / safe_n[:, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        plogp = np.where(h > 0, p * np.log2(np.where(h > 0, p, 1.0)), 0.0)
    k = (h > 0).sum(1)
    X = np.stack([-plogp.sum(1), np.log2(safe_n), ln.max(1), (h * ln).sum(1) / safe_n,
                  np.log2(np.maximum(k, 1))], 1)
    X[n <= 0] = 0.0
    return X


def features(hist, lengths):
    """Feature vector (float64[5]) of a buffer from its 256 counts and code lengths by byte:
    entropy (bits), log2 size, longest and mean code length, log2 symbol count."""
    return features_batch(np.asarray(hist)[None], np.asarray(lengths)[None])[0]


CODEBOOK_DTYPE = np.dtype(_capi.Codebook)
INFO_DTYPE = np.dtype(DectableInfo)


def codebook_array(codebooks):
    """A contiguous ctypes array of hgap_codebook (copies when given a list)."""
    if isinstance(codebooks, ctypes.Array):
        return codebooks
    arr = (_capi.Codebook * max(len(codebooks), 1))()
    for i, cb in enumerate(codebooks):
        ctypes.memmove(ctypes.byref(arr, i * ctypes.sizeof(_capi.Codebook)), ctypes.byref(cb),
                       ctypes.sizeof(_capi.Codebook))
    return arr


def code_lengths(cb_arr, count):
    """Code lengths by byte [count, 256] (uint32) of a codebook array."""
    return np.frombuffer(cb_arr, dtype=CODEBOOK_DTYPE, count=count)["length"]


def codebook_features(hist, cb):
    return features(hist, np.asarray(cb.length[:], dtype=np.float64))


def targets(enc_ms, dec_ms):
    """(target [N, N_OUT], mask [N, N_OUT]) from per-candidate times (NaN = unsupported,
    -1 = not measured)."""
    blocks_t, blocks_m = [], []
    for ms in (np.asarray(enc_ms, np.float64), np.asarray(dec_ms, np.float64)):
        measured = ms != -1.0
        ok = measured & ~np.isnan(ms) & (ms > 0)
        best = np.where(ok, ms, np.inf).min(axis=1, keepdims=True)
        t = np.where(ok, np.log2(np.where(ok, ms, 1.0) / np.where(np.isfinite(best), best, 1.0)),
                     UNSUPPORTED_PENALTY)
        blocks_t.append(t)
        blocks_m.append(measured.astype(np.float64))
    return np.concatenate(blocks_t, 1), np.concatenate(blocks_m, 1)


def regret(ms, choice):
    """Aggregate and mean per-buffer regret of choosing candidate `choice[i]` for buffer i, with
    per-candidate times `ms` [N, C] (NaN = unsupported)."""
    ms = np.asarray(ms, np.float64)
    ok = ~np.isnan(ms) & (ms > 0)
    best = np.where(ok, ms, np.inf).min(1)
    chosen = ms[np.arange(len(ms)), np.asarray(choice)]
    if np.isnan(chosen).any():
        raise ValueError("a choice is an unsupported shape")
    return dict(aggregate=float(chosen.sum() / best.sum() - 1.0),
                mean=float((chosen / best).mean() - 1.0),
                chosen_ms=float(chosen.sum()), best_ms=float(best.sum()))


# ---- network on the device -------------------------------------------------------------------

def _f32(t):
    return ctypes.c_void_p(t.data_ptr())


class Network:
    """Two-layer perceptron trained by libhgap's device kernels (fp32)."""

    def __init__(self, inputs, hidden, outputs, device="cuda", seed=0, params=None):
        import torch

        self.lib = _capi.load()
        self.dims = (int(inputs), int(hidden), int(outputs))
        self.count = int(self.lib.hgap_mlp_param_count(*self.dims))
        self.device = device
        if params is None:
            params = self.init_params(seed)
        params = np.asarray(params, np.float32)
        if params.size != self.count:
            raise ValueError(f"expected {self.count} parameters, got {params.size}")
        self.params = torch.from_numpy(params.copy()).to(device)
        self.grad = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step_count = 0

    def init_params(self, seed):
        """He-uniform first layer, small second layer, zero biases."""
        i, h, o = self.dims
        rng = np.random.default_rng(seed)
        w1 = rng.uniform(-1, 1, (h, i)) * math.sqrt(6.0 / i)
        w2 = rng.uniform(-1, 1, (o, h)) * math.sqrt(6.0 / h) * 0.1
        return np.concatenate([w1.ravel(), np.zeros(h), w2.ravel(), np.zeros(o)])

    def _stream(self, stream):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def forward(self, x, stream=None):
        """(hidden [n, H], output [n, O]) device tensors for device features x [n, I]."""
        import torch

        i, h, o = self.dims
        x = x.contiguous()
        n = x.shape[0]
        hid = torch.empty((n, h), dtype=torch.float32, device=self.device)
        y = torch.empty((n, o), dtype=torch.float32, device=self.device)
        check(self.lib.hgap_mlp_forward(_f32(x), n, i, h, o, _f32(self.params), _f32(hid),
                                        _f32(y), self._stream(stream)), "hgap_mlp_forward")
        return hid, y

    def gradient(self, x, t, m, mask_sum=None, stream=None):
        """Fills self.grad with dL/dparams at the current parameters; returns the outputs."""
        import torch

        i, h, o = self.dims
        n = x.shape[0]
        if mask_sum is None:
            mask_sum = float(m.sum().item())
        hid, y = self.forward(x, stream)
        dy = torch.empty_like(y)
        dh = torch.empty_like(hid)
        check(self.lib.hgap_mlp_backward(_f32(x), _f32(hid), _f32(y), _f32(t), _f32(m), n, i, h,
                                         o, mask_sum, _f32(self.params), _f32(dy), _f32(dh),
                                         _f32(self.grad), self._stream(stream)),
              "hgap_mlp_backward")
        return y

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        self.step_count += 1
        check(self.lib.hgap_mlp_adam_step(_f32(self.params), _f32(self.grad), _f32(self.m),
                                          _f32(self.v), self.count, self.step_count, lr, beta1,
                                          beta2, eps, self._stream(stream)),
              "hgap_mlp_adam_step")

    def fit(self, x, t, m, steps, lr, log_every=0):
        """Full-batch Adam on device tensors; returns [(step, loss)] sampled every log_every."""
        import torch

        msum = float(m.sum().item())
        hist = []
        for s in range(steps):
            y = self.gradient(x, t, m, msum)
            if log_every and (s % log_every == 0 or s == steps - 1):
                hist.append((s, float((m * (y - t) ** 2).sum().item()) / msum))
            self.adam(lr)
        torch.cuda.current_stream().synchronize()
        return hist

    def host_params(self):
        return self.params.cpu().numpy().astype(np.float32)


# Networks per predictor, trained from different seeds; their predicted costs are averaged.
# With 384 training buffers one network's held-out regret varies a lot with its seed (the
# learned excess over the exhaustive optimum ranged from 0.41 to 0.55 of the default shapes'
# excess over three seeds); the average of four ranged from 0.34 to 0.46 over two sweeps.
ENSEMBLE = 4


class Predictor:
    """Feature normalisation + the networks + the candidate lists."""

    def __init__(self, params, mean, std, lo, hi, hidden=HIDDEN, device="cuda"):
        self.mean = np.asarray(mean, np.float64)
        self.std = np.asarray(std, np.float64)
        self.lo = np.asarray(lo, np.float64)
        self.hi = np.asarray(hi, np.float64)
        self.hidden = int(hidden)
        self.device = device
        # params: [count] for one network or [k, count] for an ensemble
        self.nets = [Network(len(FEATURE_NAMES), self.hidden, N_OUT, device, params=p)
                     for p in np.atleast_2d(np.asarray(params, np.float32))]

    @classmethod
    def train(cls, X, enc_ms, dec_ms, steps=3000, lr=3e-3, hidden=HIDDEN, seed=0,
              device="cuda", log_every=0, ensemble=ENSEMBLE):
        """Fits a predictor to sweep rows X [N, 5] with candidate times: `ensemble` networks
        from seeds seed, seed + 1, ...; returns (predictor, loss history of the first)."""
        import torch

        X = np.asarray(X, np.float64)
        mean, std = X.mean(0), X.std(0)
        std = np.where(std > 1e-9, std, 1.0)
        T, M = targets(enc_ms, dec_ms)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device)  # noqa
        x, t, m = dev((X - mean) / std), dev(T), dev(M)
        params, loss = [], None
        for k in range(max(1, int(ensemble))):
            net = Network(len(FEATURE_NAMES), hidden, N_OUT, device, seed=seed + k)
            hist = net.fit(x, t, m, steps, lr, log_every if k == 0 else 0)
            loss = hist if k == 0 else loss
            params.append(net.host_params())
        p = cls(np.stack(params), mean, std, X.min(0), X.max(0), hidden, device)
        return p, loss

    def save(self, path):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        params = np.stack([net.host_params() for net in self.nets])
        np.savez(path, params=params, mean=self.mean, std=self.std, lo=self.lo,
                 hi=self.hi, hidden=np.int64(self.hidden),
                 enc_candidates=json.dumps(ENC_CANDIDATES),
                 dec_candidates=json.dumps(DEC_CANDIDATES),
                 feature_names=json.dumps(FEATURE_NAMES))

    @classmethod
    def load(cls, path=DEFAULT_MODEL, device="cuda"):
        z = np.load(path, allow_pickle=False)
        if json.loads(str(z["enc_candidates"])) != ENC_CANDIDATES or \
                json.loads(str(z["dec_candidates"])) != DEC_CANDIDATES:
            raise ValueError(f"{path}: candidate lists differ from this build's search space")
        return cls(z["params"], z["mean"], z["std"], z["lo"], z["hi"], int(z["hidden"]), device)

    def costs(self, X):
        """Predicted log2 relative cost [N, N_OUT] (host float64) of every candidate: the mean
        over the networks."""
        import torch

        X = np.atleast_2d(np.asarray(X, np.float64))
        if len(X) == 0:
            return np.zeros((0, N_OUT))
        x = torch.from_numpy(((X - self.mean) / self.std).astype(np.float32)).to(self.device)
        ys = [net.forward(x)[1].cpu().numpy().astype(np.float64) for net in self.nets]
        return np.mean(ys, axis=0)

    def in_range(self, X, margin=0.05):
        X = np.atleast_2d(np.asarray(X, np.float64))
        span = np.maximum(self.hi - self.lo, 1e-9)
        return ((X >= self.lo - margin * span) & (X <= self.hi + margin * span)).all(1)

    def choose_indices(self, X):
        """(encoder index [N], decoder index [N]) of the lowest predicted cost."""
        c = self.costs(X)
        return c[:, :N_ENC].argmin(1), c[:, N_ENC:].argmin(1)


def _supported(lib, cfg, max_len, info):
    return lib.hgap_config_supported(ctypes.byref(LaunchConfig.from_dict(cfg)), int(max_len),
                                     None if info is None else ctypes.byref(info)) == 0


def dectable_infos(cb_arr, count, prefix_bits, threads=0):
    """Decode-table infos (structured numpy [count]) for first-level widths prefix_bits[i]."""
    infos = (DectableInfo * max(count, 1))()
    pb = np.ascontiguousarray(np.broadcast_to(np.asarray(prefix_bits, np.uint32), (count,)))
    check(_capi.load().hgap_dectable_build_batch(cb_arr, count, pb.ctypes.data_as(ctypes.c_void_p),
                                                 infos, None, None, 0, threads),
          "hgap_dectable_build_batch")
    return np.frombuffer(infos, dtype=INFO_DTYPE, count=count).copy()


def support_masks(cb_arr, count, max_lens, threads=0):
    """(encoder [count, N_ENC], decoder [count, N_DEC]) boolean masks of the candidate shapes
    the device can hold for each buffer (hgap_config_supported; one query per distinct key)."""
    from . import tuner

    lib = _capi.load()
    ml = np.asarray(max_lens, np.int64)
    enc = np.zeros((count, N_ENC), bool)
    for L in np.unique(ml):
        enc[ml == L] = [_supported(lib, c, max(int(L), 1), None) for c in ENC_CANDIDATES]
    dec = np.zeros((count, N_DEC), bool)
    for p in tuner.PREFIX_BITS:
        iv = dectable_infos(cb_arr, count, p, threads)
        ms = iv["max_symbols_per_segment"].astype(np.int64)
        nbytes = iv["bytes"].astype(np.int64)
        # the check depends on the table's size and the segment density only: one query per pair
        keys, inv = np.unique(np.stack([nbytes, ms], 1), axis=0, return_inverse=True)
        cols = [j for j, c in enumerate(DEC_CANDIDATES) if c["prefix_bits"] == p]
        sup = np.zeros((len(keys), len(cols)), bool)
        for u in range(len(keys)):
            info = DectableInfo()
            info.bytes = int(keys[u, 0])
            info.max_symbols_per_segment = int(keys[u, 1])
            sup[u] = [_supported(lib, DEC_CANDIDATES[j], 0, info) for j in cols]
        dec[:, cols] = sup[inv.reshape(-1)]
    return enc, dec


def choose_indices(predictor, X, enc_mask, dec_mask):
    """Per buffer, the supported encoder and decoder candidates of lowest predicted cost."""
    c = predictor.costs(X)
    ce = np.where(enc_mask, c[:, :N_ENC], np.inf)
    cd = np.where(dec_mask, c[:, N_ENC:], np.inf)
    if not np.isfinite(ce.min(1)).all() or not np.isfinite(cd.min(1)).all():
        raise RuntimeError("no candidate launch shape fits this device for some buffer")
    return ce.argmin(1), cd.argmin(1)


def choose_configs(hists, codebooks, predictor=None, buffers=None, fallback=True, threads=0):
    """Per-buffer (encoder config, decoder config, learned mask).  With a predictor, buffers whose
    features lie in its training range get its choice among the shapes the device can hold; the
    others (or all, without a predictor) are tuned exhaustively on the device when `fallback` and
    `buffers` are given, else get the default shapes."""
    B = len(hists)
    cb_arr = codebook_array(codebooks)
    enc = [dict(DEFAULT_ENC) for _ in range(B)]
    dec = [dict(DEFAULT_DEC) for _ in range(B)]
    learned = np.zeros(B, bool)
    if predictor is not None and B:
        lengths = code_lengths(cb_arr, B)
        X = features_batch(np.asarray(hists, np.float64).reshape(B, 256), lengths)
        learned = predictor.in_range(X)
        idx = np.nonzero(learned)[0]
        if len(idx):
            em, dm = support_masks(cb_arr, B, lengths.max(1), threads)
            ei, di = choose_indices(predictor, X[idx], em[idx], dm[idx])
            for r, i in enumerate(idx):
                enc[i] = dict(ENC_CANDIDATES[ei[r]])
                dec[i] = dict(DEC_CANDIDATES[di[r]])
    if fallback and buffers is not None:
        from . import codec, tuner

        for i in np.nonzero(~learned)[0]:
            b = buffers[i]
            if b.numel() == 0:
                continue
            cb = cb_arr[i]
            sz = codec.encode_sizes(cb, hists[i])
            enc[i], _ = tuner.tune_encode(b, cb, sz, reps=3)
            e = codec.encode(b, cb=cb, sizes=sz, config=enc[i])
            dec[i], _ = tuner.tune_decode(e, reps=3)
    return enc, dec, learned


# ---- sweeps ----------------------------------------------------------------------------------

def sweep_spec(count, seed=11, min_bytes=4 << 10, max_bytes=256 << 20):
    """Buffers of a training sweep: log-uniform sizes; Zipf(s ~ U(0.6, 2.0)) for 3 of 4,
    geometric(p ~ U(0.02, 0.5)) or a spike (P(0) ~ U(0.5, 0.95)) for the rest; own permutation
    each.  Deterministic in `seed`."""
    from . import synth

    out = []
    lo, hi = math.log(min_bytes), math.log(max_bytes)
    for i in range(count):
        u = [(synth.splitmix64_scalar(seed, 5 * i + j) >> 11) / float(synth.ONE53)
             for j in range(4)]
        n = int(math.exp(lo + (hi - lo) * u[0]))
        if u[1] < 0.75:
            kind, param = "zipf", 0.6 + 1.4 * u[2]
        elif u[1] < 0.875:
            kind, param = "geometric", 0.02 + 0.48 * u[2]
        else:
            kind, param = "spike", 0.5 + 0.45 * u[2]
        out.append(dict(n=n, kind=kind, param=param, seed=(seed << 32) + i,
                        perm_seed=synth.splitmix64_scalar(seed, 5 * i + 4)))
    return out


def _pmf(kind, param):
    from . import synth

    if kind == "zipf":
        return synth.pmf_zipf(param)
    if kind == "geometric":
        return synth.pmf_geometric(param)
    return synth.pmf_skew(param)


def run_sweep(spec, reps=5, log=None):
    """Times every candidate shape on every buffer of `spec`; returns a dict of arrays."""
    import torch

    from . import codec, synth, tuner

    N = len(spec)
    X = np.zeros((N, len(FEATURE_NAMES)))
    enc_ms = np.full((N, N_ENC), np.nan)
    dec_ms = np.full((N, N_DEC), np.nan)
    cap = max(d["n"] for d in spec)
    pool = torch.empty(cap + 16, dtype=torch.uint8, device="cuda")
    t0 = time.time()
    for i, d in enumerate(spec):
        b = pool[:d["n"]]
        codec.synth_fill(b, d["seed"], synth.cdf_from_pmf(_pmf(d["kind"], d["param"])),
                         synth.permutation(d["perm_seed"]))
        hist = codec.host_histogram(codec.histogram(b))
        cb = codec.codebook_build(hist)
        sz = codec.encode_sizes(cb, hist)
        X[i] = codebook_features(hist, cb)
        _, res = tuner.tune_encode(b, cb, sz, space=ENC_CANDIDATES, reps=reps)
        enc_ms[i] = [np.nan if ms is None else ms for _, ms in res]
        e = codec.encode(b, cb=cb, sizes=sz)
        _, res = tuner.tune_decode(e, space=DEC_CANDIDATES, reps=reps)
        dec_ms[i] = [np.nan if ms is None else ms for _, ms in res]
        if log and (i % 16 == 0 or i == N - 1):
            log(f"sweep {i + 1}/{N} n={d['n']} {d['kind']} {time.time() - t0:.1f}s")
    return dict(X=X, enc_ms=enc_ms, dec_ms=dec_ms, n=np.array([d["n"] for d in spec]),
                kind=np.array([d["kind"] for d in spec]),
                param=np.array([d["param"] for d in spec]))


def load_sweeps(path=DEFAULT_SWEEPS):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


def split(sweeps, holdout=0.25):
    """(train, held-out) row indices: every 4th buffer is held out."""
    N = len(sweeps["X"])
    every = max(2, int(round(1.0 / holdout)))
    held = np.arange(N) % every == every - 1
    return np.nonzero(~held)[0], np.nonzero(held)[0]


def evaluate(pred, sweeps, rows):
    """Regret of the predictor, of the default shapes and of the best single fixed shape (chosen
    on the same rows, an optimistic baseline) against exhaustive tuning on `rows`."""
    X = sweeps["X"][rows]
    enc_ms, dec_ms = sweeps["enc_ms"][rows], sweeps["dec_ms"][rows]
    ei, di = pred.choose_indices(X)
    # a predicted shape that is unsupported for a buffer falls back to the next-best prediction
    c = pred.costs(X)
    for r in range(len(rows)):
        if np.isnan(enc_ms[r, ei[r]]):
            ei[r] = next(j for j in np.argsort(c[r, :N_ENC]) if not np.isnan(enc_ms[r, j]))
        if np.isnan(dec_ms[r, di[r]]):
            di[r] = next(j for j in np.argsort(c[r, N_ENC:]) if not np.isnan(dec_ms[r, j]))
    de = candidate_index(DEFAULT_ENC, ENC_CANDIDATES)
    dd = candidate_index(DEFAULT_DEC, DEC_CANDIDATES)
    out = {}
    for name, ms, pick, dflt in (("encode", enc_ms, ei, de), ("decode", dec_ms, di, dd)):
        fixed = np.nanargmin(np.where(np.isnan(ms).any(0), np.inf, ms.sum(0)))
        out[name] = dict(learned=regret(ms, pick),
                         default=regret(ms, np.full(len(ms), dflt)),
                         best_fixed=regret(ms, np.full(len(ms), fixed)),
                         best_fixed_shape=(ENC_CANDIDATES if name == "encode"
                                           else DEC_CANDIDATES)[fixed])
        out[name]["gain_vs_default"] = (out[name]["default"]["chosen_ms"]
                                        / out[name]["learned"]["chosen_ms"])
    out["rows"] = int(len(rows))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m huffman_coding_gap_arrays_amd.predictor")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("sweep")
    a.add_argument("--count", type=int, default=512)
    a.add_argument("--seed", type=int, default=11)
    a.add_argument("--reps", type=int, default=5)
    a.add_argument("--out", required=True)
    a = sub.add_parser("train")
    a.add_argument("--sweeps", default=DEFAULT_SWEEPS)
    a.add_argument("--out", required=True)
    a.add_argument("--steps", type=int, default=3000)
    a.add_argument("--lr", type=float, default=3e-3)
    a.add_argument("--seed", type=int, default=0)
    a.add_argument("--all", action="store_true", help="train on every row (no held-out rows)")
    a = sub.add_parser("eval")
    a.add_argument("--sweeps", default=DEFAULT_SWEEPS)
    a.add_argument("--model", default=DEFAULT_MODEL)
    args = ap.parse_args(argv)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)  # noqa: E731
    if args.cmd == "sweep":
        res = run_sweep(sweep_spec(args.count, args.seed), args.reps, log)
        np.savez(args.out, **res)
        log(f"wrote {args.out}")
        return 0
    sw = load_sweeps(args.sweeps)
    train_rows, held = split(sw)
    if args.cmd == "train":
        rows = np.arange(len(sw["X"])) if args.all else train_rows
        t0 = time.time()
        pred, loss = Predictor.train(sw["X"][rows], sw["enc_ms"][rows], sw["dec_ms"][rows],
                                     steps=args.steps, lr=args.lr, seed=args.seed,
                                     log_every=max(1, args.steps // 10))
        log(f"trained on {len(rows)} buffers in {time.time() - t0:.2f}s; loss {loss}")
        pred.save(args.out)
        print(json.dumps(dict(train_rows=len(rows), loss=loss,
                              held_out=evaluate(pred, sw, held))), flush=True)
        return 0
    pred = Predictor.load(args.model)
    print(json.dumps(evaluate(pred, sw, held)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

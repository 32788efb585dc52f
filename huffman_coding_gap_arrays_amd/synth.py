"""Synthetic inputs of the benchmark configurations.

Every buffer comes from a counter-based generator, so the CPU and the GPU produce the same
bytes: element i of stream `seed` is splitmix64 evaluated at counter i, its top 53 bits pick a
rank by integer inverse CDF (first rank whose 53-bit cumulative threshold exceeds the value),
and the byte is perm[rank].  The thresholds are computed once here, in float64, from the
closed-form probabilities of each distribution.

Configurations (sizes in bytes):
  geometric_16MiB : P(k) proportional to 0.1 * 0.9^k, k = 0..255, byte = k        (16 MiB)
  zipf_1GiB       : P(rank r) proportional to r^-1.1, r = 1..256, byte = perm(r)   (1 GiB)
  gauss_q8_512MiB : clamp(rint(8 x), -128, 127) as int8, x ~ N(0, 1)               (512 MiB)
  skew90_512MiB   : P(0) = 0.9, P(1..255) = 0.1 / 255                             (512 MiB)
  uniform_512MiB  : P(b) = 1/256                                                   (512 MiB)
"""
import math

import numpy as np

GOLDEN = 0x9E3779B97F4A7C15
ONE53 = 1 << 53
_M64 = (1 << 64) - 1


def splitmix64_at(seed, idx):
    """splitmix64 output for counters `idx` (uint64 array) of stream `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def splitmix64_scalar(seed, i):
    z = (seed + (i + 1) * GOLDEN) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def cdf_from_pmf(pmf):
    """53-bit cumulative thresholds (uint64[256]) for a 256-entry pmf; the last is 2^53."""
    p = np.asarray(pmf, dtype=np.float64)
    p = p / p.sum()
    cum = np.cumsum(p)
    t = np.floor(cum * float(ONE53)).astype(np.float64)
    t = np.minimum(t, float(ONE53)).astype(np.uint64)
    t[-1] = np.uint64(ONE53)
    t = np.maximum.accumulate(t)
    return t.astype(np.uint64)


def permutation(seed):
    """Fisher-Yates shuffle of 0..255 driven by the splitmix64 stream `seed`."""
    perm = list(range(256))
    for step, i in enumerate(range(255, 0, -1)):
        j = splitmix64_scalar(seed, step) % (i + 1)
        perm[i], perm[j] = perm[j], perm[i]
    return np.array(perm, dtype=np.uint8)


def pmf_geometric(p=0.1):
    k = np.arange(256, dtype=np.float64)
    return p * (1.0 - p) ** k


def pmf_zipf(s):
    r = np.arange(1, 257, dtype=np.float64)
    return r ** (-s)


def pmf_gauss_q8(step=1.0 / 8.0):
    """P(v) for v = clamp(rint(x / step), -128, 127), x ~ N(0,1); rank = v + 128."""
    def phi(x):
        return 0.5 * math.erfc(-x / math.sqrt(2.0))

    out = np.zeros(256)
    for r in range(256):
        v = r - 128
        lo = -math.inf if v == -128 else (v - 0.5) * step
        hi = math.inf if v == 127 else (v + 0.5) * step
        out[r] = (1.0 if hi == math.inf else phi(hi)) - (0.0 if lo == -math.inf else phi(lo))
    return out


def pmf_skew(p0=0.9):
    out = np.full(256, (1.0 - p0) / 255.0)
    out[0] = p0
    return out


def pmf_uniform():
    return np.full(256, 1.0 / 256.0)


IDENTITY = np.arange(256, dtype=np.uint8)
GAUSS_BYTES = np.array([(r - 128) & 0xFF for r in range(256)], dtype=np.uint8)

CONFIGS = {
    "geometric_16MiB": dict(n=16 << 20, seed=1, pmf=lambda: pmf_geometric(0.1), perm=IDENTITY),
    "zipf_1GiB": dict(n=1 << 30, seed=2, pmf=lambda: pmf_zipf(1.1), perm=("perm", 0x5EED2)),
    "gauss_q8_512MiB": dict(n=512 << 20, seed=3, pmf=pmf_gauss_q8, perm=GAUSS_BYTES),
    "skew90_512MiB": dict(n=512 << 20, seed=4, pmf=lambda: pmf_skew(0.9), perm=IDENTITY),
    "uniform_512MiB": dict(n=512 << 20, seed=5, pmf=pmf_uniform, perm=IDENTITY),
}


def spec(name):
    """(n, seed, cdf uint64[256], perm uint8[256]) of a named configuration."""
    c = CONFIGS[name]
    perm = c["perm"]
    if isinstance(perm, tuple):
        perm = permutation(perm[1])
    return c["n"], c["seed"], cdf_from_pmf(c["pmf"]()), np.asarray(perm, dtype=np.uint8)


def zipf_spec(s, seed, perm_seed):
    """Distribution of one Zipf(s) buffer with its own seeded rank-to-byte permutation."""
    return cdf_from_pmf(pmf_zipf(s)), permutation(perm_seed), seed


def fill_numpy(n, start, seed, cdf, perm, chunk=1 << 22):
    """numpy restatement of the generator (CPU, for checks on small sizes)."""
    out = np.empty(n, dtype=np.uint8)
    cdf = np.asarray(cdf, dtype=np.uint64)
    perm = np.asarray(perm, dtype=np.uint8)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        r = splitmix64_at(seed, np.arange(start + a, start + b, dtype=np.uint64)) >> np.uint64(11)
        out[a:b] = perm[np.searchsorted(cdf, r, side="right")]
    return out


def batch_spec(count=4096, seed=4, min_bytes=4 << 10, max_bytes=64 << 20, s_lo=0.6, s_hi=2.0):
    """Buffers of the batched workload: sizes log-uniform in [min_bytes, max_bytes], each with
    its own Zipf exponent ~ U(s_lo, s_hi) and its own seeded rank-to-byte permutation.
    Returns a list of dicts (n, s, seed, perm_seed); deterministic in `seed`."""
    out = []
    lo, hi = math.log(min_bytes), math.log(max_bytes)
    for i in range(count):
        u1 = (splitmix64_scalar(seed, 3 * i) >> 11) / float(ONE53)
        u2 = (splitmix64_scalar(seed, 3 * i + 1) >> 11) / float(ONE53)
        n = int(math.exp(lo + (hi - lo) * u1))
        out.append(dict(n=n, s=s_lo + (s_hi - s_lo) * u2, seed=(seed << 32) + i,
                        perm_seed=splitmix64_scalar(seed, 3 * i + 2)))
    return out

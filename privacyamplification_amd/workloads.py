"""Synthetic workloads of the measurement plan: block shapes and generator seeds only.

Every block is (n, m, key_seed, toeplitz_seed): the key is synth_bits(key_seed, n + 1) and the
seed synth_bits(toeplitz_seed, n) (xoshiro256** through splitmix64, Bernoulli(0.5)), XOR key rest
and compression on, as the reference's key and seed servers send them (examples/*).  bench.py
times these workloads; tests/golden/make_manifest.py writes the exact outputs' SHA3-256 digests
for them, so every output of a timed run can be checked.
"""
import numpy as np


def batched_mix(count=256, master=0xB47C):
    """Uneven blocks: n = 32*floor(2^U/32), U ~ Uniform[16, 25]; m/n uniform in {1/8, 1/4, 3/8,
    1/2}, m a multiple of 32; key / seed generator seeds 0x5000+2i / 0x5001+2i."""
    rng = np.random.default_rng(master)
    blocks = []
    for i in range(count):
        n = int(2 ** rng.uniform(16, 25)) // 32 * 32
        m = max(32, int(n * rng.choice([1, 2, 3, 4]) / 8) // 32 * 32)
        blocks.append((n, m, 0x5000 + 2 * i, 0x5001 + 2 * i))
    return blocks


SERVE_STREAMS = 4       # static seed streams, each with a resident seed
SERVE_KEYS = 8          # distinct keys per stream (cycled)
SERVE_DYNAMIC = 48      # distinct dynamic-seed blocks
SERVE_STATIC_SHARE = 0.75


def serving_set(master=0x5E4E):
    """The blocks a serving stream draws from.  Returns (streams, dynamic):
    streams[k] = (n, m, toeplitz_seed, [key_seed, ...]): a shape from the batched mix, its resident
    seed 0x5EED+k and keys 0xC000+16k+j; dynamic[j] = (n, m, key_seed, toeplitz_seed), the first
    SERVE_DYNAMIC blocks of the batched mix drawn with this master seed."""
    rng = np.random.default_rng(master)
    mix = batched_mix(count=256, master=master)
    streams = []
    for k in range(SERVE_STREAMS):
        n, m, _, _ = mix[int(rng.integers(0, len(mix)))]
        streams.append((n, m, 0x5EED + k, [0xC000 + 16 * k + j for j in range(SERVE_KEYS)]))
    return streams, list(mix[:SERVE_DYNAMIC])


def serving_arrivals(rate_gbit_s, count, mean_n, master=0x5E4E):
    """Poisson arrivals (seconds from t = 0) at rate_gbit_s of offered key bits, and per block the
    draw that picks it: (times, kinds, picks); kinds < SERVE_STATIC_SHARE selects a static stream."""
    rng = np.random.default_rng((master, int(rate_gbit_s * 1000) if np.isfinite(rate_gbit_s) else 0))
    if np.isfinite(rate_gbit_s):
        times = np.cumsum(rng.exponential(mean_n / (rate_gbit_s * 1e9), count))
    else:
        times = np.zeros(count)
    return times, rng.random(count), rng.integers(0, 1 << 30, count)


def serving_pick(streams, dynamic, kind, pick):
    """The block a draw selects: ('static', stream index, key index) or ('dynamic', index)."""
    if kind < SERVE_STATIC_SHARE:
        k = int(pick % len(streams))
        return ("static", k, int(pick // len(streams) % len(streams[k][3])))
    return ("dynamic", int(pick % len(dynamic)))


def serving_mean_n(streams, dynamic):
    return (SERVE_STATIC_SHARE * float(np.mean([s[0] for s in streams]))
            + (1 - SERVE_STATIC_SHARE) * float(np.mean([d[0] for d in dynamic])))

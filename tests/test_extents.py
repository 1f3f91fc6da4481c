"""Buffer extents (pa_block_extents, host only): for every block shape the engine accepts, no kernel
indexes past any buffer that an allocation path gives it, and the caller-visible windows are the
reference's own sizes: the key (n/32+1 words with the key-rest XOR, h/32+1 without,
PrivacyAmplification.cu:1428, 1068-1072), the seed (n/32 words) and the output (m/8 bytes with
compression, n/16 without, :1186), the last checked against the oracle's output length."""
import numpy as np
import pytest

import privacyamplification_amd as P

FLAG_PAIRS = [(True, True), (False, True), (True, False), (False, False)]


def shapes():
    rng = np.random.default_rng(0xE7)
    out = []
    for e in range(6, 30):
        for n in {1 << e, (1 << e) + 32, 3 << max(e - 2, 4), (1 << e) - 32}:
            if n < 64 or n > (1 << 29):
                continue
            ms = {32, n - 32, n // 8 // 32 * 32, 3 * n // 8 // 32 * 32, n // 2 // 32 * 32,
                  7 * n // 8 // 32 * 32, int(rng.integers(1, n // 32)) * 32}
            out += [(n, m) for m in sorted(ms) if 0 < m < n]
    return out


SHAPES = shapes()


def test_sweep_covers_the_size_range():
    ns = {n for n, _ in SHAPES}
    assert min(ns) == 64 and max(ns) == 1 << 29 and len(SHAPES) > 500


@pytest.mark.parametrize("do_xor,do_compress", FLAG_PAIRS)
def test_no_kernel_indexes_past_an_allocation(do_xor, do_compress):
    flags = (P.XOR_KEY_REST if do_xor else 0) | (P.COMPRESS if do_compress else 0)
    checked = 0
    for n, m in SHAPES:
        pow2 = n & (n - 1) == 0
        if not pow2 and not do_compress:
            with pytest.raises(P.EngineError):
                P.block_extents(n, m, flags)
            continue
        L, touched, alloc = P.block_extents(n, m, flags)
        assert L >= n and L & (L - 1) == 0 and L < 2 * n
        for b in P.BUFFERS:
            assert touched[b] <= alloc[b], (n, m, flags, b, touched[b], alloc[b])
        h = n - m
        assert touched["key"] == 4 * (n // 32 + 1 if do_xor else h // 32 + 1)
        assert touched["seed"] == n // 8
        assert touched["out"] == P.out_bytes(n, m, do_compress)
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("n,m", [(64, 32), (2048, 1024), (4096, 1536), (16384, 6144), (65536, 57344), (96000, 24000)])
@pytest.mark.parametrize("do_xor,do_compress", FLAG_PAIRS)
def test_output_window_is_the_oracle_output_length(oracle, n, m, do_xor, do_compress):
    if n & (n - 1) and not do_compress:
        pytest.skip("n/2-bit output is defined for power-of-two n only")
    flags = (P.XOR_KEY_REST if do_xor else 0) | (P.COMPRESS if do_compress else 0)
    _, touched, _ = P.block_extents(n, m, flags)
    key = oracle.synth_bits(7, n + 1)
    if not do_xor:
        key = key[: (n - m) // 32 + 1].copy()
    ref = oracle.amp_out(key, oracle.synth_bits(8, n), n, m, do_xor, do_compress, method=oracle.DIRECT)
    assert touched["out"] == ref.size

"""The oracle against everything that pins it: the reference's golden vectors (LargeBlocksizeExample,
Matrix.h), its unit tests' closed forms (PrivacyAmplification.cu:299-737), brute force, and GF(2)
invariants.  CPU only."""
import hashlib

import numpy as np
import pytest

from oracle import lbs

METHODS = ("DIRECT", "NTT", "FFT32", "FFT64")


def sha3(b):
    return hashlib.sha3_256(bytes(b)).hexdigest()


@pytest.mark.parametrize("method", METHODS)
def test_lbs_tiles_and_running_states(oracle, lbs_fixture, method):
    """G1/G2: 60 tile outputs, 70 running XOR states, pre-XOR, key-rest and final digests
    (LargeBlocksizeExample/main.cpp:437-523 checks exactly these)."""
    m = getattr(oracle, method)
    tile = lambda k, s: oracle.amp_out(k, s, 2048, 1024, do_xor=False, do_compress=False, method=m)
    chunks, inter, pre, kr, res, _ = lbs.run_schedule(lbs_fixture, tile)
    assert chunks == lbs_fixture["chunk_sha3"]
    assert inter == lbs_fixture["intermediate_sha3"]
    assert pre == lbs_fixture["pre_xor_sha3"]
    assert kr == lbs_fixture["key_rest_sha3"]
    assert res == lbs_fixture["result_sha3"]


@pytest.mark.parametrize("method", METHODS)
def test_lbs_as_one_xor_block(oracle, lbs_fixture, method):
    """G3: the whole tiling equals one XOR-mode block at n=2^14, m=6144 with key 0||K and the
    seed rotated left by h; this pins key2StartRest's key-rest extraction and the XOR path."""
    key, seed, n, m = lbs.full_block_inputs(lbs_fixture)
    out = oracle.amp_out(key, seed, n, m, method=getattr(oracle, method))
    assert sha3(out) == lbs_fixture["result_sha3"]


def _brute(key_start, kbits, seed, n, nout):
    x = [(key_start[j >> 5] >> (31 - (j & 31))) & 1 for j in range(kbits)]
    s = [(seed[j >> 5] >> (31 - (j & 31))) & 1 for j in range(n)]
    return [sum(x[j] & s[(p - j) % n] for j in range(kbits)) & 1 for p in range(nout)]


def _bits_msb(b, nbits):
    return [(b[p >> 3] >> (7 - (p & 7))) & 1 for p in range(nbits)]


@pytest.mark.parametrize("n,kbits,nout", [(64, 40, 24), (128, 97, 32), (256, 161, 96), (512, 257, 256)])
@pytest.mark.parametrize("method", METHODS)
def test_conv_against_brute_force(oracle, n, kbits, nout, method):
    rng = np.random.default_rng(n + kbits)
    ks = rng.integers(0, 2**32, n // 32, dtype=np.uint64).astype(np.uint32)
    seed = rng.integers(0, 2**32, n // 32, dtype=np.uint64).astype(np.uint32)
    if method in ("FFT32", "FFT64") and n < 128:
        pytest.skip("FFT restatement needs n >= 128")
    got = oracle.conv_bits(ks, kbits, seed, n, nout, getattr(oracle, method))
    assert _bits_msb(got, nout) == _brute(ks, kbits, seed, n, nout)


@pytest.mark.parametrize("n,kbits,nout", [(96, 40, 56), (160, 97, 63), (480, 200, 280), (992, 700, 292)])
def test_ntt_any_length_against_brute_force(oracle, n, kbits, nout):
    """Non-power-of-two n: the NTT's linear-convolution form equals brute force mod n."""
    rng = np.random.default_rng(n * 7 + kbits)
    ks = rng.integers(0, 2**32, n // 32, dtype=np.uint64).astype(np.uint32)
    seed = rng.integers(0, 2**32, n // 32, dtype=np.uint64).astype(np.uint32)
    got = oracle.conv_bits(ks, kbits, seed, n, nout, oracle.NTT)
    assert _bits_msb(got, nout) == _brute(ks, kbits, seed, n, nout)


@pytest.mark.parametrize("n,m", [(96000, 24000), (100000 - 100000 % 32, 25024), (5 << 12, 1 << 12)])
def test_ntt_any_length_matches_direct(oracle, n, m):
    key = oracle.synth_bits(n, n + 1)
    seed = oracle.synth_bits(n + 1, n)
    for do_xor in (True, False):
        k = key if do_xor else key[: (n - m) // 32 + 1].copy()
        assert np.array_equal(oracle.amp_out(k, seed, n, m, do_xor, True, method=oracle.DIRECT),
                              oracle.amp_out(k, seed, n, m, do_xor, True, method=oracle.NTT))


@pytest.mark.parametrize("n,kbits,nout", [(256, 161, 96), (480, 200, 280), (1024, 700, 512)])
def test_long_length_ntt_against_brute_force(oracle, n, kbits, nout):
    """The NTT over 3*2^30+1 (used for transforms longer than 2^27) forced on small cases."""
    rng = np.random.default_rng(n * 11 + kbits)
    ks = rng.integers(0, 2**32, n // 32, dtype=np.uint64).astype(np.uint32)
    seed = rng.integers(0, 2**32, n // 32, dtype=np.uint64).astype(np.uint32)
    oracle.set_ntt_big(True)
    try:
        got = oracle.conv_bits(ks, kbits, seed, n, nout, oracle.NTT)
        counts = oracle.conv_counts(ks, kbits, seed, n, nout) if n & (n - 1) == 0 else None
    finally:
        oracle.set_ntt_big(False)
    assert _bits_msb(got, nout) == _brute(ks, kbits, seed, n, nout)
    if counts is not None:
        assert np.array_equal(counts, oracle.conv_counts(ks, kbits, seed, n, nout))


def test_long_length_ntt_matches_direct(oracle):
    n, m = 1 << 16, 3 << 13
    key = oracle.synth_bits(0x31, n + 1)
    seed = oracle.synth_bits(0x32, n)
    oracle.set_ntt_big(True)
    try:
        got = oracle.amp_out(key, seed, n, m, method=oracle.NTT)
    finally:
        oracle.set_ntt_big(False)
    assert np.array_equal(got, oracle.amp_out(key, seed, n, m, method=oracle.DIRECT))


@pytest.mark.parametrize("e", [10, 12, 16, 20])
def test_methods_agree(oracle, e):
    n = 1 << e
    m = 3 * n // 8
    key = oracle.synth_bits(0x11 + e, n + 1)
    seed = oracle.synth_bits(0x22 + e, n)
    outs = {meth: oracle.amp_out(key, seed, n, m, method=getattr(oracle, meth)).tobytes()
            for meth in METHODS}
    assert len(set(outs.values())) == 1


@pytest.mark.parametrize("n,m", [(512, 384), (1024, 992), (1024, 256)])
@pytest.mark.parametrize("method", METHODS)
def test_xor_without_compression_against_brute_force(oracle, n, m, method):
    """XOR mode without compression sends n/2 output bits (sendData, PrivacyAmplification.cu:
    1186-1197); ToBinaryArray XORs the key rest into each of them that has one (:753-777): all
    n/2 when m > n/2, the first m otherwise."""
    h = n - m
    key = oracle.synth_bits(0x5A + n + m, n + 1)
    seed = oracle.synth_bits(0x5B + n + m, n)
    got = _bits_msb(oracle.amp_out(key, seed, n, m, do_xor=True, do_compress=False,
                                   method=getattr(oracle, method)), n // 2)
    conv = _brute(key, h + 1, seed, n, n // 2)
    rest = [(key[(h + 1 + p) >> 5] >> (31 - ((h + 1 + p) & 31))) & 1 if p < m else 0
            for p in range(n // 2)]
    assert got == [a ^ b for a, b in zip(conv, rest)]


def test_ntt_counts_exact(oracle):
    """The NTT returns the exact integer convolution (not only its parity)."""
    n, kbits = 1024, 641
    rng = np.random.default_rng(5)
    ks = rng.integers(0, 2**32, n // 32, dtype=np.uint64).astype(np.uint32)
    seed = rng.integers(0, 2**32, n // 32, dtype=np.uint64).astype(np.uint32)
    x = np.array([(ks[j >> 5] >> (31 - (j & 31))) & 1 for j in range(kbits)], dtype=np.int64)
    s = np.array([(seed[j >> 5] >> (31 - (j & 31))) & 1 for j in range(n)], dtype=np.int64)
    ref = np.array([sum(int(x[j]) * int(s[(p - j) % n]) for j in range(kbits)) for p in range(n)])
    got = oracle.conv_counts(ks, kbits, seed, n, n)
    assert np.array_equal(got.astype(np.int64), ref)


def _flip(words, bit):
    w = words.copy()
    w[bit >> 5] ^= np.uint32(1 << (31 - (bit & 31)))
    return w


def test_gf2_invariants(oracle):
    n, m = 4096, 1536
    h = n - m
    key = oracle.synth_bits(7, n + 1)
    seed = oracle.synth_bits(8, n)
    zero_key = np.zeros_like(key)
    # zero key -> zero output; linearity in the key (the map is linear, key rest included)
    assert not oracle.amp_out(zero_key, seed, n, m).any()
    k2 = oracle.synth_bits(9, n + 1)
    a, b = oracle.amp_out(key, seed, n, m), oracle.amp_out(k2, seed, n, m)
    assert np.array_equal(oracle.amp_out(key ^ k2, seed, n, m), a ^ b)
    # a key with only bit j (j <= h) set gives column j of T: y[p] = s[(p-j) mod n]
    for j in (0, 1, 77, h):
        e = _flip(np.zeros_like(key), j)
        y = _bits_msb(oracle.amp_out(e, seed, n, m), m)
        s = [(seed[i >> 5] >> (31 - (i & 31))) & 1 for i in range(n)]
        assert y == [s[(p - j) % n] for p in range(m)]
    # key bits above h only reach the key rest: y = r, shifted by one (key2StartRest)
    only_rest = key.copy()
    for j in range(h + 1):
        if (only_rest[j >> 5] >> (31 - (j & 31))) & 1:
            only_rest = _flip(only_rest, j)
    y = _bits_msb(oracle.amp_out(only_rest, seed, n, m), m)
    assert y == [(key[(h + 1 + p) >> 5] >> (31 - ((h + 1 + p) & 31))) & 1 for p in range(m)]
    # seed e_0: T is the identity on the first m key bits
    e0 = _flip(np.zeros(n // 32, dtype=np.uint32), 0)
    y = _bits_msb(oracle.amp_out(key, e0, n, m, do_xor=True), m)
    x = [(key[p >> 5] >> (31 - (p & 31))) & 1 for p in range(m)]
    r = [(key[(h + 1 + p) >> 5] >> (31 - ((h + 1 + p) & 31))) & 1 for p in range(m)]
    assert y == [a_ ^ b_ for a_, b_ in zip(x, r)]


def _a000788(n):
    """A000788: total number of 1 bits in 0..n (PrivacyAmplification.cu:484-491 uses it)."""
    return sum(bin(i).count("1") for i in range(n + 1))


@pytest.mark.parametrize("e", [10, 12, 14])
def test_bin_int2float_closed_form(oracle, e):
    """U4 (PrivacyAmplification.cu:525-568): binIn[i] = i; each float is 0 or 2^-11 per the
    MSB-first bit; the popcount is A000788(n/32 - 1)."""
    n = 1 << e
    bits = np.arange(n // 32, dtype=np.uint32)
    f, ones = oracle.bin_int2float(bits, n, 11)
    assert ones == _a000788(n // 32 - 1)
    idx = np.arange(n)
    expect = np.where(((idx // 32) & (1 << (31 - (idx % 32)))) != 0, np.float32(1 / 2048), 0)
    assert np.array_equal(f, expect.astype(np.float32))


def test_correction_closed_form(oracle):
    """U1 (PrivacyAmplification.cu:299-361) and the TEST-mode constant 0x3F54D912 (:2024)."""
    for i in range(64):
        for j in range(64):
            got = oracle.correction(i, j, 64)
            assert abs(got - ((i * j) / 64.0) % 2.0) < 1e-4
    rng = np.random.default_rng(777)
    for _ in range(4096):
        a, b = (int(v) for v in rng.integers(2**25, 2**27, 2))
        ref = (((a * b) & 0xFFFFFFFF) / float(2**27)) % 2.0
        assert abs(oracle.correction(a, b, 2**27) - ref) < 1e-4
    v = np.float32(oracle.correction(41947248, 67113455, 2**27))
    assert v.view(np.uint32) == 0x3F54D912


def test_element_wise_product_closed_form(oracle):
    """U3 (PrivacyAmplification.cu:422-463): do1[i]=i+0.77, do2[i]=i+0.88, r=32."""
    cnt = 1024
    a = np.arange(2 * cnt, dtype=np.float32) + np.float32(0.77)
    b = np.arange(2 * cnt, dtype=np.float32) + np.float32(0.88)
    got = oracle.element_wise_product(a, b, 32.0)
    ar, ai = a[0::2].astype(np.float64) / 32, a[1::2].astype(np.float64) / 32
    br, bi = b[0::2].astype(np.float64) / 32, b[1::2].astype(np.float64) / 32
    assert np.allclose(got[0::2], ar * br - ai * bi, rtol=1e-6, atol=1e-3)
    assert np.allclose(got[1::2], ar * bi + ai * br, rtol=1e-6, atol=1e-3)


@pytest.mark.parametrize("e", [10, 13, 16])
@pytest.mark.parametrize("do_xor", [True, False])
def test_to_binary_array_closed_form(oracle, e, do_xor):
    """U5 (PrivacyAmplification.cu:605-737): invOut[p] = bit p of (p/32) as 0/1, corr = 1.9,
    norm = 1, key_rest = 0xAA..; output bit p = bit ^ key_rest bit in the OutBytes layout."""
    n = 1 << e
    m = n // 4 + n // 8
    idx = np.arange(n)
    v = np.where(((idx // 32) & (1 << (31 - (idx % 32)))) != 0, 1.0, 0.0).astype(np.float32)
    kr = np.full(n // 32, 0xAAAAAAAA, dtype=np.uint32)
    words = oracle.to_binary_array(v, m, 1.0, 1.9, kr if do_xor else None, True)
    out = words.view(np.uint8)
    for p in range(0, m, 97):
        bit = (out[p >> 3] >> (7 - (p & 7))) & 1
        expect = ((p // 32) >> (31 - (p % 32))) & 1  # rint(x + 1.9) is odd iff x is 1... (x+2 -> even)
        expect ^= 0  # rint(0 + 1.9) = 2 (even), rint(1 + 1.9) = 3 (odd): the bit survives
        if do_xor:
            expect ^= (0xAAAAAAAA >> (31 - (p % 32))) & 1
        assert bit == expect


def test_reverse_endian_false_is_refbits(oracle):
    n, m = 2048, 768
    key = oracle.synth_bits(3, n + 1)
    seed = oracle.synth_bits(4, n)
    a = oracle.amp_out(key, seed, n, m, reverse_endian=True)
    b = oracle.amp_out(key, seed, n, m, reverse_endian=False)
    assert np.array_equal(b.view("<u4").byteswap().view(np.uint8), a)


def test_synth_generator_known_values(oracle):
    """xoshiro256** from splitmix64(0): first output pinned from the published reference
    implementation's algorithm (state after seeding is splitmix64's first four outputs)."""
    w = oracle.synth_bits(0, 128)
    # splitmix64 stream for seed 0 starts 0xe220a8397b1dcdaf; xoshiro256** output = rotl(s1*5,7)*9
    sm = [0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f, 0xf88bb8a8724c81ec]
    mask = (1 << 64) - 1
    r = ((((sm[1] * 5) & mask) << 7 | ((sm[1] * 5) & mask) >> 57) & mask) * 9 & mask
    assert (int(w[0]) << 32 | int(w[1])) == r

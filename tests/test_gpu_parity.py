"""Bit-exact parity of the HIP engine with the oracle and with the reference's golden vectors.

Runs on an MI355X (marker gpu).  Every case prints its parameters before the GPU call."""
import hashlib
import json
import pathlib

import numpy as np
import pytest

from oracle import lbs

pytestmark = pytest.mark.gpu
ROOT = pathlib.Path(__file__).resolve().parents[1]


def sha3(b):
    return hashlib.sha3_256(bytes(b)).hexdigest()


def say(*a):
    print("[case]", *a, flush=True)


def test_lbs_golden_tiles(engine, lbs_fixture):
    """G1/G2: the 60 tiles of LargeBlocksizeExample through the engine (n=2^11, m=1024, NoXOR,
    no compression) reproduce every digest of Matrix.h."""
    say("lbs tiles n=2048 noxor nocompress")
    tile = lambda k, s: engine.amplify(k, s, 2048, 1024, do_xor_key_rest=False, do_compress=False)
    chunks, inter, pre, kr, res, _ = lbs.run_schedule(lbs_fixture, tile)
    assert chunks == lbs_fixture["chunk_sha3"]
    assert inter == lbs_fixture["intermediate_sha3"]
    assert (pre, kr, res) == (lbs_fixture["pre_xor_sha3"], lbs_fixture["key_rest_sha3"],
                              lbs_fixture["result_sha3"])


def test_lbs_golden_single_block(engine, lbs_fixture):
    say("lbs single xor block n=16384")
    key, seed, n, m = lbs.full_block_inputs(lbs_fixture)
    assert sha3(engine.amplify(key, seed, n, m)) == lbs_fixture["result_sha3"]


CASES = [(e, frac) for e in range(10, 21) for frac in ((3, 8),)] + \
        [(e, frac) for e in (11, 14, 15, 17) for frac in ((1, 8), (1, 4), (1, 2))]


@pytest.mark.parametrize("e,frac", CASES)
@pytest.mark.parametrize("do_xor,do_compress", [(True, True), (False, True), (False, False), (True, False)])
def test_random_blocks_vs_oracle(engine, oracle, e, frac, do_xor, do_compress):
    n = 1 << e
    m = n * frac[0] // frac[1]
    say(f"random n=2^{e} m={m} xor={do_xor} compress={do_compress}")
    key = oracle.synth_bits(0x100 + e, n + 1)
    seed = oracle.synth_bits(0x200 + e, n)
    if not do_xor:
        key = key[: (n - m) // 32 + 1].copy()
    ref = oracle.amp_out(key, seed, n, m, do_xor, do_compress, method=oracle.NTT)
    got = engine.amplify(key, seed, n, m, do_xor, do_compress)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("e,frac,do_xor,do_compress", [(24, (3, 8), False, False), (27, (3, 8), True, False),
                                                       (27, (1, 8), True, True), (26, (1, 2), False, True),
                                                       (25, (1, 4), True, False)])
def test_large_blocks_flags_and_ratios(engine, oracle, e, frac, do_xor, do_compress):
    """The output windows of the other flags and m/n ratios at full transform sizes (pass A'
    covers n/2 bits without compression, m bits with it), against the NTT oracle."""
    n = 1 << e
    m = n * frac[0] // frac[1]
    say(f"large n=2^{e} m={m} xor={do_xor} compress={do_compress}")
    key = oracle.synth_bits(0x300 + e, n + 1)
    seed = oracle.synth_bits(0x400 + e, n)
    if not do_xor:
        key = key[: (n - m) // 32 + 1].copy()
    ref = oracle.amp_out(key, seed, n, m, do_xor, do_compress, method=oracle.NTT)
    assert np.array_equal(engine.amplify(key, seed, n, m, do_xor, do_compress), ref)


@pytest.mark.parametrize("e", [12, 14, 16, 20, 22])
def test_static_seed_matches_dynamic(engine, oracle, e):
    n, m = 1 << e, 3 * (1 << e) // 8
    say(f"static seed n=2^{e}")
    seed = oracle.synth_bits(0x5EED, n)
    sid = engine.seed_put(seed, n, m)
    try:
        for i in range(3):
            key = oracle.synth_bits(0x1000 + i, n + 1)
            ref = oracle.amp_out(key, seed, n, m, method=oracle.NTT)
            assert np.array_equal(engine.amplify_seeded(key, sid, n, m), ref)
            assert np.array_equal(engine.amplify(key, seed, n, m), ref)
    finally:
        engine.seed_drop(sid)


@pytest.mark.parametrize("n,m", [(3 << 21, 3 << 18), (20229632, 5057408)])
def test_static_seed_non_power_of_two(engine, oracle, n, m):
    """A resident seed for a non-2^k block: its spectrum is built from the rotated seed repeated up
    to L, like the dynamic path's, and serves several keys."""
    say(f"static seed non-2^k n={n} m={m}")
    seed = oracle.synth_bits(0x5EEE, n)
    sid = engine.seed_put(seed, n, m)
    try:
        for i in range(2):
            key = oracle.synth_bits(0x1100 + i, n + 1)
            ref = oracle.amp_out(key, seed, n, m, method=oracle.NTT)
            assert np.array_equal(engine.amplify_seeded(key, sid, n, m), ref)
    finally:
        engine.seed_drop(sid)


@pytest.mark.parametrize("p_key", [0.9, 0.1, 1.0, 0.0])
def test_biased_keys(engine, oracle, p_key):
    n, m = 1 << 18, 3 << 15
    say(f"biased key p={p_key} n=2^18")
    key = oracle.synth_bits(0x77, n + 1, p_key)
    seed = oracle.synth_bits(0x78, n)
    ref = oracle.amp_out(key, seed, n, m, method=oracle.NTT)
    assert np.array_equal(engine.amplify(key, seed, n, m), ref)


@pytest.mark.parametrize("n,m", [(3 << 15, 3 << 13), (5 << 12, 1 << 12), (100000 - 100000 % 32, 25024)])
def test_non_power_of_two_blocks(engine, oracle, n, m):
    """Blocks of n = 32k bits run on a power-of-two transform with the seed rotated by m
    (same Toeplitz product y[p] = XOR_j x[j] s[(p-j) mod n])."""
    say(f"non-2^k n={n} m={m}")
    key = oracle.synth_bits(n, n + 1)
    seed = oracle.synth_bits(n + 1, n)
    ref = oracle.amp_out(key, seed, n, m, method=oracle.DIRECT)
    assert np.array_equal(engine.amplify(key, seed, n, m), ref)


@pytest.mark.parametrize("n,m", [((1 << 23) + 32, 1 << 20), (8664960, 1083104), (10852704, 2713152),
                                 (20229632, 5057408), (26088768, 6522176), ((1 << 25) - 32, 3 << 22),
                                 ((1 << 26) + 32, 1 << 24), (3 << 25, 3 << 23)])
def test_large_non_power_of_two_blocks(engine, oracle, n, m):
    """Non-2^k blocks whose transform length L = 2^ceil(log2 n) leaves up to half of L beyond n:
    the engine repeats the rotated seed up to L, so the values before rounding stay small (a zero
    pad there makes them swing by about h/4 and FP32 rounding fails from L = 2^24 on).  Checked
    against the exact NTT oracle, one block at a time and in one batch, with the rounding margin.
    The last two run on the 2^27 transform, whose pass B writes tiles for pass A', with output
    windows that start inside the transform."""
    import torch
    say(f"large non-2^k n={n} m={m}")
    key = oracle.synth_bits(0xA00 + n % 997, n + 1)
    seed = oracle.synth_bits(0xB00 + n % 991, n)
    ref = oracle.amp_out(key, seed, n, m, method=oracle.NTT)
    dev = torch.device("cuda:0")
    kd = torch.from_numpy(np.concatenate([key, np.zeros(1, np.uint32)]).view(np.int32)).to(dev)
    sd = torch.from_numpy(seed.view(np.int32)).to(dev)
    out = torch.zeros(m // 8 + 16, dtype=torch.uint8, device=dev)
    margin = torch.zeros(1, dtype=torch.float32, device=dev)
    engine.block_device(kd.data_ptr(), sd.data_ptr(), n, m, 3, out.data_ptr(), 0, margin.data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(out[: m // 8].cpu().numpy(), ref)
    assert float(margin.item()) < 0.25, float(margin.item())
    outs = engine.run_batch([{"n": n, "m": m, "key": key, "seed": seed},
                             {"n": n, "m": m, "key": key, "seed": seed}])
    assert all(np.array_equal(o, ref) for o in outs)


@pytest.mark.parametrize("e,m_of", [(16, lambda n: n - 32), (20, lambda n: 7 * n // 8), (23, lambda n: 7 * n // 8),
                                    (23, lambda n: n - 32)])
def test_output_longer_than_half_block(engine, oracle, e, m_of):
    """m > n/2: the compressed output (m/8 bytes) is longer than the n/16 bytes of the
    uncompressed one, and the host-buffer path's staging must hold it."""
    n = 1 << e
    m = m_of(n)
    say(f"long output n=2^{e} m={m}")
    key = oracle.synth_bits(0xE00 + e, n + 1)
    seed = oracle.synth_bits(0xF00 + e, n)
    ref = oracle.amp_out(key, seed, n, m, method=oracle.NTT)
    assert np.array_equal(engine.amplify(key, seed, n, m), ref)
    ticket, out, keep = engine.submit(key, seed, n, m)  # asynchronous path, its own staging
    engine.wait(ticket)
    assert np.array_equal(out[: m // 8], ref)
    # without compression only n/2 bits go out, each with its key-rest bit XORed in
    ref_nc = oracle.amp_out(key, seed, n, m, True, False, method=oracle.NTT)
    assert np.array_equal(engine.amplify(key, seed, n, m, True, False), ref_nc)


def _edge_cases():
    out = []
    for n in (64, 96, 128, 160, 192, 256, 288, 512, 544, 16384, 16416, 32800):
        pow2 = n & (n - 1) == 0
        for m in sorted({32, n // 2 // 32 * 32, n - 32}):
            for do_xor in (True, False):
                for do_compress in ((True, False) if pow2 else (True,)):
                    out.append((n, m, do_xor, do_compress))
    return out


@pytest.mark.parametrize("n,m,do_xor,do_compress", _edge_cases())
def test_smallest_blocks_and_path_boundaries(engine, oracle, n, m, do_xor, do_compress):
    """The smallest n (64) and m (32), m = n - 32, non-2^k sizes just above a power of two, and the
    sizes either side of the one-workgroup path's limit (L = 2^14 / 2^15)."""
    say(f"edge n={n} m={m} xor={do_xor} compress={do_compress}")
    key = oracle.synth_bits(0x1E00 + n + m, n + 1)
    if not do_xor:
        key = key[: (n - m) // 32 + 1].copy()
    seed = oracle.synth_bits(0x1F00 + n + m, n)
    ref = oracle.amp_out(key, seed, n, m, do_xor, do_compress, method=oracle.NTT)
    assert np.array_equal(engine.amplify(key, seed, n, m, do_xor, do_compress), ref)


@pytest.mark.parametrize("m_of", [lambda n: 32, lambda n: n - 32], ids=["m=32", "m=n-32"])
def test_largest_block_extreme_ratios(engine, oracle, m_of):
    """n = 2^27 with the shortest and the longest output."""
    n = 1 << 27
    m = m_of(n)
    say(f"extreme ratio n=2^27 m={m}")
    key = oracle.synth_bits(0x2E00 + m % 997, n + 1)
    seed = oracle.synth_bits(0x2F00 + m % 997, n)
    ref = oracle.amp_out(key, seed, n, m, method=oracle.NTT)
    assert np.array_equal(engine.amplify(key, seed, n, m), ref)


def test_random_sweep_vs_oracle(engine, oracle):
    """40 seeded random blocks across the whole size range, against the NTT oracle:
    * n log-uniform in 2^10..2^24, half of them non-2^k (compressed, as those require);
    * m any multiple of 32 below n;
    * random XOR and compress flags where defined, and some biased keys.
    Each block runs alone and, all together, in one batch."""
    rng = np.random.default_rng(0x5EE9)
    blocks, refs = [], []
    for i in range(40):
        e = rng.uniform(10, 24)
        pow2 = bool(rng.random() < 0.5)
        n = 1 << int(round(e)) if pow2 else max(64, int(2 ** e) // 32 * 32)
        m = int(rng.integers(1, n // 32)) * 32
        do_xor = bool(rng.random() < 0.7)
        do_compress = True if not pow2 else bool(rng.random() < 0.7)
        p_key = float(rng.choice([0.5, 0.5, 0.9, 0.1]))
        key = oracle.synth_bits(0xC00 + i, n + 1, p_key)
        if not do_xor:
            key = key[: (n - m) // 32 + 1].copy()
        seed = oracle.synth_bits(0xD00 + i, n)
        say(f"sweep {i}: n={n} m={m} xor={do_xor} compress={do_compress} p={p_key}")
        ref = oracle.amp_out(key, seed, n, m, do_xor, do_compress, method=oracle.NTT)
        assert np.array_equal(engine.amplify(key, seed, n, m, do_xor, do_compress), ref), i
        blocks.append({"n": n, "m": m, "key": key, "seed": seed, "do_xor_key_rest": do_xor,
                       "do_compress": do_compress})
        refs.append(ref)
    outs = engine.run_batch(blocks)
    for i, (o, ref) in enumerate(zip(outs, refs)):
        assert np.array_equal(o, ref), i


def test_async_submit_wait(engine, oracle):
    n, m = 1 << 16, 3 << 13
    say("async submit x4")
    jobs = []
    for i in range(4):
        key = oracle.synth_bits(0x900 + i, n + 1)
        seed = oracle.synth_bits(0xA00 + i, n)
        t, out, keep = engine.submit(key, seed, n, m)
        jobs.append((t, out, keep, oracle.amp_out(key, seed, n, m, method=oracle.NTT)))
    for t, out, keep, ref in jobs:
        engine.wait(t)
        assert np.array_equal(out[: ref.size], ref)


def test_refbits_output_order(oracle):
    import privacyamplification_amd as P
    n, m = 1 << 15, 3 << 12
    say("reverse_endian=0 engine")
    key = oracle.synth_bits(1, n + 1)
    seed = oracle.synth_bits(2, n)
    with P.Engine(out_reverse_endian=False) as eng:
        got = eng.amplify(key, seed, n, m)
    assert np.array_equal(got, oracle.amp_out(key, seed, n, m, reverse_endian=False, method=oracle.NTT))


def _manifest():
    return json.loads((ROOT / "tests/golden/manifest.json").read_text())["blocks"]


@pytest.fixture(scope="module")
def big_engine():
    """An engine accepting transforms beyond the reference's 2^27 (up to 2^29)."""
    import privacyamplification_amd as P
    eng = P.Engine(max_n_log2=29)
    yield eng
    eng.close()


@pytest.mark.parametrize("item", _manifest(), ids=lambda it: it["name"])
def test_golden_manifest(engine, big_engine, oracle, item):
    """Full-size blocks (2^20 .. 2^29) against the committed oracle digests; 2^28 and 2^29 exceed the
    reference's largest block (oracle-only parity)."""
    n, m = item["n"], item["m"]
    say(f"manifest {item['name']} n={n}")
    key = oracle.synth_bits(item["key_seed"], n + 1, item["p_key"])
    seed = oracle.synth_bits(item["toeplitz_seed"], n)
    eng = engine if n <= 1 << 27 else big_engine
    assert sha3(eng.amplify(key, seed, n, m)) == item["sha3"]


def test_golden_manifest_batched(engine, oracle):
    """The batched kernels at the reference's largest size: the four 2^27 manifest blocks (two with
    their own seeds, two on one resident seed) in one batch, against the committed digests."""
    items = [it for it in _manifest() if it["n"] == 1 << 27]
    assert len(items) == 4
    say("manifest 2^27 blocks in one batch")
    blocks, sids = [], []
    try:
        for it in items:
            n, m = it["n"], it["m"]
            key = oracle.synth_bits(it["key_seed"], n + 1, it["p_key"])
            seed = oracle.synth_bits(it["toeplitz_seed"], n)
            if "static" in it["name"]:
                if not sids:  # both static blocks use the same seed: one resident spectrum
                    sids.append(engine.seed_put(seed, n, m))
                blocks.append({"n": n, "m": m, "key": key, "seed": None, "sid": sids[0]})
            else:
                blocks.append({"n": n, "m": m, "key": key, "seed": seed})
        outs = engine.run_batch(blocks)
        for it, o in zip(items, outs):
            assert sha3(o) == it["sha3"], it["name"]
    finally:
        for sid in sids:
            engine.seed_drop(sid)


def test_golden_manifest_batched_large(big_engine, oracle):
    """The batched kernels beyond the reference's largest size: a 2^28 and a 2^29 manifest block in
    one batch (pass B's tiles and the paired-column pass A' at N1 = 2^14, the single-column one at
    2^15), against the committed digests."""
    items = [it for it in _manifest() if it["name"] in ("c3_n28_p0.5", "c3_n29_p0.9")]
    assert len(items) == 2
    say("manifest 2^28 + 2^29 blocks in one batch")
    blocks = []
    for it in items:
        n, m = it["n"], it["m"]
        key = oracle.synth_bits(it["key_seed"], n + 1, it["p_key"])
        seed = oracle.synth_bits(it["toeplitz_seed"], n)
        blocks.append({"n": n, "m": m, "key": key, "seed": seed})
    outs = big_engine.run_batch(blocks)
    for it, o in zip(items, outs):
        assert sha3(o) == it["sha3"], it["name"]


def test_linearity_at_full_size(engine, oracle):
    """amp_out is GF(2)-linear in the key (T x and the key rest both are): out(a^b) = out(a)^out(b)."""
    n, m = 1 << 27, 3 << 24
    say("linearity n=2^27")
    seed = oracle.synth_bits(0x2000, n)
    a = oracle.synth_bits(0xA, n + 1)
    b = oracle.synth_bits(0xB, n + 1)
    oa, ob = engine.amplify(a, seed, n, m), engine.amplify(b, seed, n, m)
    assert np.array_equal(engine.amplify(a ^ b, seed, n, m), oa ^ ob)


CANARY_SHAPES = [(1 << 10, 1 << 9, 3), (1 << 14, 6144, 3), (1 << 15, 7 << 12, 2), (1 << 15, 3 << 12, 0),
                 (1 << 20, (1 << 20) - 32, 2), (1 << 20, 3 << 17, 1), (96000, 24000, 3), (3 << 20, 7 << 18, 2),
                 (1 << 24, 32, 3), (1 << 24, 7 << 21, 0), ((1 << 25) + 32, 5 << 20, 2)]


@pytest.mark.parametrize("n,m,flags", CANARY_SHAPES)
def test_output_writes_stay_in_window(engine, oracle, n, m, flags):
    """The output stage writes exactly the pa_block_extents window (m/8 bytes with compression,
    n/16 without): a canary past it survives, and the window equals the oracle."""
    import torch
    import privacyamplification_amd as P
    do_xor, do_compress = bool(flags & 1), bool(flags & 2)
    _, touched, _ = P.block_extents(n, m, flags)
    nb = touched["out"]
    assert nb == P.out_bytes(n, m, do_compress)
    say(f"canary n={n} m={m} flags={flags}")
    key = oracle.synth_bits(0x600 + n % 977, n + 1)
    seed = oracle.synth_bits(0x700 + n % 977, n)
    if not do_xor:
        key = key[: (n - m) // 32 + 1].copy()
    dev = torch.device("cuda:0")
    dk = torch.from_numpy(np.concatenate([key, np.zeros(2, np.uint32)]).view(np.int32)).to(dev)
    ds = torch.from_numpy(seed.view(np.int32)).to(dev)
    out = torch.full((nb + 512,), 0xA5, dtype=torch.uint8, device=dev)
    engine.block_device(dk.data_ptr(), ds.data_ptr(), n, m, flags, out.data_ptr())
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.all(got[nb:] == 0xA5), np.nonzero(got[nb:] != 0xA5)[0][:8]
    assert np.array_equal(got[:nb], oracle.amp_out(key, seed, n, m, do_xor, do_compress, method=oracle.NTT))

"""Batched launches (pa_run_batch / pa_batch_device): many blocks of uneven length, flags and seed
modes in one call must give, block for block, the same amp_out as the oracle and as the one-block
entry points.  Runs on an MI355X (marker gpu)."""
import numpy as np
import pytest

import privacyamplification_amd as P

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _mixed_blocks(oracle, seed_ids, rng, count, lo=10, hi=20, pow2_only=False):
    """Random blocks: n log-uniform in [2^lo, 2^hi] (multiples of 32), m/n in {1/8,1/4,3/8,1/2}
    rounded to a multiple of 32, random flags (non-2^k n needs compression), a quarter of them on
    a resident seed."""
    blocks = []
    for i in range(count):
        e = rng.uniform(lo, hi)
        n = 1 << int(round(e)) if pow2_only or rng.random() < 0.5 else int(2 ** e) // 32 * 32
        frac = rng.choice([1, 2, 3, 4]) / 8
        m = max(32, int(n * frac) // 32 * 32)
        if m >= n:
            m = n - 32
        pow2 = n & (n - 1) == 0
        do_xor = bool(rng.random() < 0.7)
        do_compress = True if not pow2 else bool(rng.random() < 0.8)
        b = {"n": n, "m": m, "do_xor_key_rest": do_xor, "do_compress": do_compress}
        key = oracle.synth_bits(0x5000 + 2 * i, n + 1)
        if not do_xor:
            key = key[: (n - m) // 32 + 1].copy()
        b["key"] = key
        if (n, m) in seed_ids and rng.random() < 0.5:
            b["sid"], b["seed_arr"] = seed_ids[(n, m)]
        else:
            b["seed"] = oracle.synth_bits(0x5001 + 2 * i, n)
            b["seed_arr"] = b["seed"]
        blocks.append(b)
    return blocks


def _ref(oracle, b):
    return oracle.amp_out(b["key"], b["seed_arr"], b["n"], b["m"], b["do_xor_key_rest"], b["do_compress"],
                          method=oracle.NTT if b["n"] & (b["n"] - 1) == 0 else oracle.DIRECT)


@pytest.mark.parametrize("bflags", [0, P.BATCH_GRAPH, P.BATCH_TUNE | P.BATCH_GRAPH])
def test_mixed_batch_vs_oracle(engine, oracle, bflags):
    rng = np.random.default_rng(0xB47C + bflags)
    # resident seeds for two shapes that the batch will reuse
    seed_ids = {}
    for n, m in ((1 << 14, 3 << 11), (1 << 16, 1 << 14)):
        s = oracle.synth_bits(0x5EED + n, n)
        seed_ids[(n, m)] = (engine.seed_put(s, n, m), s)
    try:
        blocks = _mixed_blocks(oracle, seed_ids, rng, 24, 10, 18)
        for n, m in seed_ids:  # make sure both resident shapes occur
            sid, s = seed_ids[(n, m)]
            blocks.append({"n": n, "m": m, "do_xor_key_rest": True, "do_compress": True,
                           "key": oracle.synth_bits(n + 7, n + 1), "sid": sid, "seed_arr": s})
        say(f"mixed batch of {len(blocks)} flags={bflags}")
        for rep in range(2):  # the second call replays the captured graph / tuned shape
            outs = engine.run_batch(blocks, bflags)
            for b, o in zip(blocks, outs):
                assert np.array_equal(o, _ref(oracle, b)), (rep, b["n"], b["m"], b["do_xor_key_rest"],
                                                           b["do_compress"], "sid" in b)
    finally:
        for sid, _ in seed_ids.values():
            engine.seed_drop(sid)


def test_batch_equals_single_blocks_large(engine, oracle):
    """Blocks up to 2^24 (all four-step column widths) in one batch equal the one-block path."""
    rng = np.random.default_rng(7)
    blocks = []
    for i, e in enumerate((15, 17, 19, 21, 22, 23, 24, 21, 19)):
        n = 1 << e if i % 2 == 0 else (1 << e) - 32 * (1 + i)
        m = (n // 8 * int(rng.integers(1, 5))) // 32 * 32
        blocks.append({"n": n, "m": m, "key": oracle.synth_bits(0x600 + i, n + 1),
                       "seed": oracle.synth_bits(0x700 + i, n)})
    say(f"large batch {[b['n'] for b in blocks]}")
    outs = engine.run_batch(blocks, P.BATCH_GRAPH)
    for b, o in zip(blocks, outs):
        assert np.array_equal(o, engine.amplify(b["key"], b["seed"], b["n"], b["m"])), b["n"]


def test_batch_device_with_margin(engine, oracle):
    import torch
    dev = torch.device("cuda:0")
    say("batch_device torch buffers")
    shapes = [(1 << 16, 3 << 13), (1 << 18, 1 << 15), (96000, 24000), (1 << 12, 1 << 9)]
    keep, arr, refs = [], (P.PaBlock * len(shapes))(), []
    margins = torch.zeros(len(shapes), dtype=torch.float32, device=dev)
    for i, (n, m) in enumerate(shapes):
        key = oracle.synth_bits(0x800 + i, n + 1)
        seed = oracle.synth_bits(0x900 + i, n)
        kd = torch.from_numpy(np.concatenate([key, np.zeros(1, np.uint32)]).view(np.int32)).to(dev)
        sd = torch.from_numpy(seed.view(np.int32)).to(dev)
        od = torch.zeros(m // 8 + 16, dtype=torch.uint8, device=dev)
        keep += [kd, sd, od]
        arr[i].key, arr[i].seed, arr[i].n, arr[i].m, arr[i].flags = kd.data_ptr(), sd.data_ptr(), n, m, 3
        arr[i].out = od.data_ptr()
        arr[i].margin = margins.data_ptr() + 4 * i
        refs.append((od, oracle.amp_out(key, seed, n, m, method=oracle.NTT if n & (n - 1) == 0 else oracle.DIRECT)))
    s = torch.cuda.Stream(device=dev)
    engine.batch_device(arr, P.BATCH_GRAPH, s.cuda_stream)
    s.synchronize()
    for od, ref in refs:
        assert np.array_equal(od[: ref.size].cpu().numpy(), ref)
    mg = margins.cpu().numpy()
    assert np.all(mg > 0) and np.all(mg < 0.25), mg


def test_batch_rejects_bad_blocks(engine, oracle):
    n, m = 1 << 12, 1 << 9
    key = oracle.synth_bits(1, n + 1)
    with pytest.raises(P.EngineError):
        engine.run_batch([{"n": n, "m": m, "key": key, "sid": 12345}])  # unknown resident seed
    with pytest.raises(P.EngineError):
        engine.run_batch([{"n": n + 1, "m": m, "key": key, "seed": key}])  # n not a multiple of 32

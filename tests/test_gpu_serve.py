"""Serving (pa_serve_*): blocks submitted while earlier ones are on the GPU are admitted into
batches as slots and HBM budget free up; every result must equal the block computed alone and the
oracle.  Runs on an MI355X (marker gpu)."""
import hashlib
import json
import pathlib
import time

import numpy as np
import pytest

import privacyamplification_amd as P

pytestmark = pytest.mark.gpu
MANIFEST = json.loads((pathlib.Path(__file__).resolve().parent / "golden" / "manifest.json").read_text())["blocks"]


def say(*a):
    print("[case]", *a, flush=True)


def _ref(oracle, key, seed, n, m, do_xor=True, do_compress=True):
    return oracle.amp_out(key, seed, n, m, do_xor, do_compress, method=oracle.NTT)


def test_serve_mixed_stream_equals_alone(engine, oracle):
    rng = np.random.default_rng(0x5E4E)
    shapes_static = [(1 << 14, 3 << 11), (96000, 24000)]
    seeds = {}
    for n, m in shapes_static:
        s = oracle.synth_bits(0x5EED + n, n)
        seeds[(n, m)] = (engine.seed_put(s, n, m), s)
    engine.serve_start(slots=3, max_batch=32, hbm_budget_bytes=2 << 30)
    try:
        jobs = []
        say("serve 80 mixed blocks")
        for i in range(80):
            if rng.random() < 0.6:
                n, m = shapes_static[int(rng.integers(0, 2))]
                sid, s = seeds[(n, m)]
                key = oracle.synth_bits(0x7000 + i, n + 1)
                out = np.zeros(m // 8 + 8, np.uint8)
                t = engine.serve_submit(key, n, m, out, sid=sid)
                jobs.append((t, out, key, s, n, m, True, True))
            else:
                e = rng.uniform(11, 20)
                n = int(2 ** e) // 32 * 32
                m = max(32, int(n * rng.choice([1, 2, 3, 4]) / 8) // 32 * 32)
                pow2 = n & (n - 1) == 0
                do_xor = bool(rng.random() < 0.7)
                do_compress = True if not pow2 else bool(rng.random() < 0.8)
                key = oracle.synth_bits(0x7000 + i, n + 1)
                if not do_xor:
                    key = key[: (n - m) // 32 + 1].copy()
                s = oracle.synth_bits(0x8000 + i, n)
                out = np.zeros(P.out_bytes(n, m, do_compress) + 8, np.uint8)
                t = engine.serve_submit(key, n, m, out, seed=s, do_xor_key_rest=do_xor, do_compress=do_compress)
                jobs.append((t, out, key, s, n, m, do_xor, do_compress))
            if rng.random() < 0.3:
                time.sleep(float(rng.uniform(0, 0.002)))  # arrivals while batches are in flight
        lats = []
        for t, out, key, s, n, m, do_xor, do_compress in jobs:
            lats.append(engine.serve_wait(t, 60 * 10**9))
            nb = P.out_bytes(n, m, do_compress)
            alone = engine.amplify(key, s, n, m, do_xor, do_compress)
            assert np.array_equal(out[:nb], alone), (n, m, do_xor, do_compress)
            assert np.array_equal(out[:nb], _ref(oracle, key, s, n, m, do_xor, do_compress)), (n, m)
        assert all(x > 0 for x in lats)
    finally:
        engine.serve_stop()
        for sid, _ in seeds.values():
            engine.seed_drop(sid)


def test_serve_pinned_buffers(engine, oracle):
    """Pinned (pa_host_alloc) keys, seeds and outputs are read and written in place by the copy
    kernel, without staging; results unchanged, also when pinned and pageable blocks share a
    batch."""
    engine.serve_start(slots=2, max_batch=16)
    try:
        say("serve pinned + pageable")
        jobs = []
        for i in range(24):
            n = [1 << 16, 3 << 14, 1 << 18][i % 3]
            m = n // 4
            key = oracle.synth_bits(0xA000 + i, n + 1)
            seed = oracle.synth_bits(0xA100 + i, n)
            if i % 2 == 0:
                key, seed = engine.pinned_copy(key), engine.pinned_copy(seed)
                out = engine.host_empty(m // 8)
            else:
                out = np.zeros(m // 8, np.uint8)
            jobs.append((engine.serve_submit(key, n, m, out, seed=seed), out, key, seed, n, m))
        for t, out, key, seed, n, m in jobs:
            engine.serve_wait(t, 60 * 10**9)
            assert np.array_equal(out, _ref(oracle, np.asarray(key), np.asarray(seed), n, m)), (n, m)
    finally:
        engine.serve_stop()


def test_serve_submit_many_any_alignment(engine, oracle):
    """pa_serve_submit_many queues blocks in one call; keys, seeds and outputs at 16-byte, 4-byte
    and odd-byte offsets (pinned and pageable) all take the copy kernel's matching path."""
    n0 = 1 << 16
    seed0 = oracle.synth_bits(0xB0, n0)
    sid = engine.seed_put(seed0, n0, n0 // 4)
    engine.serve_start(slots=2, max_batch=64)
    try:
        say("serve_submit_many: 30 blocks, mixed alignment")
        reqs, meta = [], []
        for i in range(30):
            static = i % 3 == 0
            n = n0 if static else [1 << 15, 40000, 1 << 17, 100000][i % 4]
            m = n // 4 // 32 * 32
            key = oracle.synth_bits(0xB100 + i, n + 1)
            seed = seed0 if static else oracle.synth_bits(0xB200 + i, n)
            kw, ob = key.size, m // 8
            shift_k, shift_o = [(0, 0), (1, 4), (1, 1), (0, 3)][i % 4]  # words / bytes
            pinned = i % 2 == 0
            kb = engine.host_empty(kw + 2, np.uint32) if pinned else np.zeros(kw + 2, np.uint32)
            kb[shift_k:shift_k + kw] = key
            obuf = engine.host_empty(ob + 8) if pinned else np.zeros(ob + 8, np.uint8)
            out = obuf[shift_o:shift_o + ob]
            sd = None
            if not static:
                sb = engine.host_empty(seed.size + 1, np.uint32) if pinned else np.zeros(seed.size + 1, np.uint32)
                sb[shift_k:shift_k + seed.size] = seed
                sd = sb[shift_k:shift_k + seed.size]
            reqs.append((kb[shift_k:shift_k + kw], n, m, out, sd, 0 if sd is not None else sid))
            meta.append((out, key, seed, n, m, kb, obuf))
        tickets = engine.serve_submit_many(reqs)
        assert len(set(tickets)) == len(tickets)
        for t, (out, key, seed, n, m, _, _) in zip(tickets, meta):
            engine.serve_wait(t, 60 * 10**9)
            assert np.array_equal(out, _ref(oracle, key, seed, n, m)), (n, m)
    finally:
        engine.serve_stop()
        engine.seed_drop(sid)


def test_serve_large_blocks_by_dma(engine, oracle):
    """Keys, seeds and outputs of 1 MiB or more go by DMA instead of the copy kernel; mixed with
    small ones, pinned and pageable, in one stream of batches."""
    cases = [(1 << 23, 3 << 20, True), (1 << 23, 7 << 20, False), (1 << 25, 7 << 22, True),
             (1 << 25, 3 << 22, False), (1 << 16, 1 << 14, True), (1 << 23, 3 << 20, False)]
    n0, m0 = 1 << 23, 3 << 20
    seed0 = oracle.synth_bits(0xD0, n0)
    sid = engine.seed_put(seed0, n0, m0)
    engine.serve_start(slots=3, max_batch=4, hbm_budget_bytes=48 << 30)
    try:
        say("serve large blocks by DMA")
        reqs, meta = [], []
        for i, (n, m, pinned) in enumerate(cases):
            key = oracle.synth_bits(0xD100 + i, n + 1)
            resident = i == 5
            seed = seed0 if resident else oracle.synth_bits(0xD200 + i, n)
            k = engine.pinned_copy(key) if pinned else key
            sd = None if resident else (engine.pinned_copy(seed) if pinned else seed)
            out = engine.host_empty(m // 8) if pinned else np.zeros(m // 8, np.uint8)
            reqs.append((k, n, m, out, sd, sid if resident else 0))
            meta.append((out, key, seed, n, m))
        tickets = engine.serve_submit_many(reqs)
        digests = {(b["n"], b["m"], b["key_seed"]): b["sha3"] for b in MANIFEST if b["name"].startswith("dma_")}
        for i, (t, (out, key, seed, n, m)) in enumerate(zip(tickets, meta)):
            engine.serve_wait(t, 120 * 10**9)
            if n <= 1 << 23:
                assert np.array_equal(out, _ref(oracle, key, seed, n, m)), (n, m)
            else:  # 2^25-bit blocks: the oracle's digests, committed (tests/golden/make_manifest.py)
                assert hashlib.sha3_256(out.tobytes()).hexdigest() == digests[(n, m, 0xD100 + i)], (n, m)
    finally:
        engine.serve_stop()
        engine.seed_drop(sid)


def test_serve_bad_blocks_fail_alone(engine, oracle):
    """Blocks whose sizes the engine refuses (m not a multiple of 32, m >= n, a non-2^k n without
    compression, an unknown resident seed) fail with their own status inside a bulk submission;
    the valid blocks around them complete."""
    engine.serve_start(slots=2, max_batch=16)
    try:
        say("serve bulk with refused blocks")
        n, m = 1 << 18, 3 << 15
        good, reqs = [], []
        for i in range(8):
            key = oracle.synth_bits(0xE100 + i, n + 1)
            seed = oracle.synth_bits(0xE200 + i, n)
            out = np.zeros(m // 8 + 8, np.uint8)
            reqs.append((key, n, m, out, seed, 0))
            good.append((out, key, seed))
        bad = [(oracle.synth_bits(1, n + 1), n, m + 16, np.zeros(m // 8 + 8, np.uint8), oracle.synth_bits(2, n), 0),
               (oracle.synth_bits(3, n + 1), n, n, np.zeros(n // 8 + 8, np.uint8), oracle.synth_bits(4, n), 0),
               (oracle.synth_bits(5, 100001), 100000 // 32 * 32, 3200, np.zeros(4096, np.uint8), None, 987654)]
        order = [0, 1, 100, 2, 3, 101, 4, 5, 102, 6, 7]
        allreq = [reqs[i] if i < 100 else bad[i - 100] for i in order]
        tickets = engine.serve_submit_many(allreq)
        for t, i in zip(tickets, order):
            if i < 100:
                engine.serve_wait(t, 60 * 10**9)
                out, key, seed = good[i]
                assert np.array_equal(out[: m // 8], _ref(oracle, key, seed, n, m)), i
            else:
                with pytest.raises(P.EngineError) as ei:
                    engine.serve_wait(t, 60 * 10**9)
                assert ei.value.status < 0, i
    finally:
        engine.serve_stop()


def test_serve_budget_admission(engine, oracle):
    """A small budget splits the stream into small batches; a block larger than a slot's share of
    the budget is refused with PA_ERR_NOMEM and the others still complete."""
    engine.serve_start(slots=2, max_batch=64, hbm_budget_bytes=64 << 20)
    try:
        n, m = 1 << 20, 3 << 17
        say("serve budget 64 MiB, 12 blocks of 2^20 + one 2^24")
        jobs = []
        for i in range(12):
            key = oracle.synth_bits(0x9000 + i, n + 1)
            s = oracle.synth_bits(0x9100 + i, n)
            out = np.zeros(m // 8 + 8, np.uint8)
            jobs.append((engine.serve_submit(key, n, m, out, seed=s), out, key, s))
        big_n = 1 << 24
        big_key = oracle.synth_bits(1, big_n + 1)
        big_seed = oracle.synth_bits(2, big_n)
        big_out = np.zeros(big_n // 8, np.uint8)
        tb = engine.serve_submit(big_key, big_n, 3 * big_n // 8, big_out, seed=big_seed)
        for t, out, key, s in jobs:
            engine.serve_wait(t, 60 * 10**9)
            assert np.array_equal(out[: m // 8], _ref(oracle, key, s, n, m))
        with pytest.raises(P.EngineError) as ei:
            engine.serve_wait(tb, 60 * 10**9)
        assert ei.value.status == -4
    finally:
        engine.serve_stop()


def test_destroy_while_serving_drains_first(oracle):
    """pa_destroy with blocks still queued stops the scheduler first: every queued and in-flight
    block completes against the resident seeds, workspaces and tables before they are freed."""
    eng = P.Engine()
    n0, m0 = 1 << 18, 3 << 15
    sid = eng.seed_put(oracle.synth_bits(0xDE0, n0), n0, m0)
    eng.serve_start(slots=3, max_batch=8, hbm_budget_bytes=1 << 30)
    jobs = []
    say("destroy with 60 blocks queued")
    for i in range(60):
        if i % 3 == 0:
            n, m, seed = n0, m0, None
        else:
            n = [1 << 16, 96000, 1 << 20][i % 3]
            m = n // 4 // 32 * 32
            seed = oracle.synth_bits(0xDE00 + i, n)
        key = oracle.synth_bits(0xDF00 + i, n + 1)
        out = np.zeros(m // 8 + 8, np.uint8)
        eng.serve_submit(key, n, m, out, seed=seed, sid=0 if seed is not None else sid)
        jobs.append((out, key, seed, n, m))
    eng.close()  # no serve_stop, no waits: pa_destroy must drain the queue itself
    s0 = oracle.synth_bits(0xDE0, n0)
    for i, (out, key, seed, n, m) in enumerate(jobs):
        ref = _ref(oracle, key, s0 if seed is None else seed, n, m)
        assert np.array_equal(out[: m // 8], ref), i

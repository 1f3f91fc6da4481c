#!/usr/bin/env python3
"""Throughput of the Toeplitz-hashing hot path on one MI355X.

Metric: amplified key Gbit/s at 2^27-bit blocks (Gbit/s of input key, the reference's own
convention, PrivacyAmplification.cu:1213), bit-exact; and p99 block latency in serving.

The default run prints ONE JSON line.  Its headline fields time one block per step of n = 2^27
bits, m = 3n/8, XOR and compression on, with a fresh (dynamic) Toeplitz seed per block -- the
reference's secure default (config.yaml:17) -- on synthetic Bernoulli(0.5) inputs already resident
in HBM.  The same line carries:
  static        the same block with a resident seed (the reference's reuse_seed_amount != 0 mode)
  batched       256 uneven blocks per step (n log-uniform 2^16..2^25, m/n in {1/8, 1/4, 3/8, 1/2})
                through pa_batch_device, beside one launch sequence per block on the same blocks
  serving       Poisson arrivals of that length mix at 32 Gbit/s offered load, 4 resident seed
                streams plus 25 % dynamic seeds, 64 GB budget: sustained Gbit/s, p50/p99/p99.9
  cpu_baseline  the oracle's FP32 restatement of the reference pipeline on the host cores
Every output of every part is checked against the committed oracle digests
(tests/golden/manifest.json); a mismatch sets the headline value to null and exits with status 1.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
                  [--workload all|block|batched|serving|precision] [--mode dynamic|static] [--log2n 27]
"""
import argparse
import hashlib
import json
import math
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def algorithmic_bytes(n, m, mode):
    """Compulsory HBM bytes per block of the fused three-pass design (two passes per transform over
    the n/2-complex FP32 array; unpack, product and epilogue fused; static seeds read a resident
    spectrum)."""
    if n <= (1 << 14):
        return n / 4 + m / 4
    if mode == "static":
        return 20 * n + n / 8 + m / 4
    return 24 * n + n / 4 + m / 4


def kernel_bytes(n, m):
    """Algorithmic bytes per launch of each kernel as built: the compulsory HBM reads and writes of
    its inputs and outputs. A complex FP32 spectrum of an n-bit operand is 4n bytes, its packed
    bits n/8 bytes, the output nibbles m/4 bytes."""
    return {
        "k_transpose_bits(key)": n / 8 + n / 8,
        "k_transpose_bits(key+seed)": 2 * (n / 8 + n / 8),
        "k_fwd_cols(key)": n / 8 + 4 * n,
        "k_fwd_cols(key+seed)": 2 * (n / 8 + 4 * n),
        "k_rows(seedprep)": 8 * n,
        "k_rows": 12 * n,
        "k_inv_cols": 4 * n + m / 4,
        "k_pack": m / 4 + m / 8 + m / 8,
        "k_zero_u32": 8,
    }


def read_traffic(kernel, log2n, mode):
    """HBM bytes per launch of `kernel` from the committed counter summary: a line
    `traffic <kernel> n=2^<log2n> <mode> <bytes>` (FETCH_SIZE + WRITE_SIZE, calibrated on a plain
    copy kernel)."""
    f = ROOT / "profiles" / "hbm_traffic.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        parts = line.split()
        if (len(parts) == 5 and parts[0] == "traffic" and parts[1] == kernel
                and parts[2] == f"n=2^{log2n}" and parts[3] == mode):
            return float(parts[4])
    return None


def manifest():
    """{name: entry} of the committed oracle digests (tests/golden/make_manifest.py)."""
    blocks = json.loads((ROOT / "tests/golden/manifest.json").read_text())["blocks"]
    return {b["name"]: b for b in blocks}


def sha3(b):
    return hashlib.sha3_256(bytes(b)).hexdigest()


def progress(**kw):
    print(json.dumps({"progress": kw}), file=sys.stderr, flush=True)


# ------------------------------------------------------------------ CPU baseline ----
def cpu_baseline(n, m, min_s=10.0, legs=True):
    """The reference's FP32 method restated on the host cores (oracle/, OpenMP on every core),
    timed on bounded samples of the same workloads.  Main value: n-bit blocks with fresh keys and
    seeds until at least min_s seconds of CPU work.  Legs: the PR1 size (2^20) with fresh and with
    resident seeds, 2^27 with a resident seed (its spectrum computed once, off the clock, as the
    reference does for a reused seed), and the exact NTT oracle at 2^20 and 2^24."""
    from oracle import oracle as O
    O.build()
    threads = O.get_threads()

    def leg(log2n, mode, min_secs, max_blocks=64):
        nn = 1 << log2n
        mm = 3 * nn // 8
        spec = O.fft32_seed_spectrum(O.synth_bits(0x5EED, nn), nn) if mode == "static" else None
        blocks, dt = 0, 0.0
        while blocks < max_blocks and (blocks == 0 or dt < min_secs):
            key = O.synth_bits(0x1000 + blocks, nn + 1)
            if mode == "static":
                t0 = time.perf_counter()
                O.amp_out_fft32_seeded(key, spec, nn, mm)
            else:
                seed = O.synth_bits(0x2000 + blocks, nn)
                t0 = time.perf_counter()
                O.amp_out(key, seed, nn, mm, method=O.NTT if mode == "ntt" else O.FFT32)
            dt += time.perf_counter() - t0
            blocks += 1
        return {"n": f"2^{log2n}", "mode": mode, "value": round(nn * blocks / dt / 1e9, 6), "unit": "Gbit/s",
                "blocks": blocks, "seconds": round(dt, 2)}

    log2n = n.bit_length() - 1
    main = leg(log2n, "dynamic", min_s)
    out = {"value": main["value"], "unit": "Gbit/s", "cores": threads, "kind": "port",
           "sample": f"{main['blocks']} blocks n=2^{log2n} m=3n/8 dynamic seed, FP32 FFT restatement of the "
                     f"reference pipeline (oracle/pa_oracle.c, OpenMP, {threads} threads), {main['seconds']} s"}
    if legs:
        out["legs"] = [leg(20, "dynamic", 2.0), leg(20, "static", 2.0), leg(27, "static", 4.0, 4),
                       leg(20, "ntt", 2.0), leg(24, "ntt", 2.0, 4)]
        out["legs_note"] = ("dynamic/static: FP32 restatement (static: seed spectrum reused); ntt: the exact "
                            "integer-NTT oracle (mod 15*2^27+1), m = 3n/8, XOR + compress")
    return out


# ------------------------------------------------------------------ device context ----
class Ctx:
    def __init__(self, max_n_log2=27):
        import torch
        import privacyamplification_amd as P
        self.torch, self.P = torch, P
        self.dev = torch.device("cuda:0")
        self.eng = P.Engine(device=0, max_n_log2=max_n_log2)
        # a dedicated stream: the engine treats a null stream handle as its own internal stream
        self.stream = torch.cuda.Stream(device=self.dev)
        torch.cuda.set_stream(self.stream)
        self.sptr = self.stream.cuda_stream

    def to_dev(self, words, pad_words=0):
        import numpy as np
        a = words if not pad_words else np.concatenate([words, np.zeros(pad_words, np.uint32)])
        return self.torch.from_numpy(a.view(np.int32)).to(self.dev)


def run_single(ctx, n, m, mode, steps, warmup, man, with_kernels=True):
    """One block per step, 2 distinct resident input sets alternated (keys 0x1000+i; seeds 0x2000+i,
    or the resident seed 0x5EED).  Both inputs' outputs are checked against the committed digests
    before the timed region, and the output of the last timed block after it."""
    import numpy as np
    torch, P, eng, stream, sptr = ctx.torch, ctx.P, ctx.eng, ctx.stream, ctx.sptr
    flags = P.DEFAULT_FLAGS
    nblk = 2
    keys = [ctx.to_dev(P.synth_bits(0x1000 + i, n + 1)) for i in range(nblk)]
    seeds = [ctx.to_dev(P.synth_bits(0x2000 + i, n)) for i in range(nblk)] if mode == "dynamic" else []
    out = torch.empty(m // 8 + 16, dtype=torch.uint8, device=ctx.dev)
    margin = torch.zeros(1, dtype=torch.float32, device=ctx.dev)
    sid = eng.seed_put(P.synth_bits(0x5EED, n), n, m) if mode == "static" else 0

    def step(i, mg=None):
        j = i % nblk
        mp = mg.data_ptr() if mg is not None else None
        if mode == "static":
            eng.block_device_seeded(keys[j].data_ptr(), sid, n, m, flags, out.data_ptr(), sptr, mp)
        else:
            eng.block_device(keys[j].data_ptr(), seeds[j].data_ptr(), n, m, flags, out.data_ptr(), sptr, mp)

    def digest_of(j):
        e = man.get(f"c2_{mode}_{j}")
        return e["sha3"] if e and e["n"] == n and e["m"] == m else None

    verified = []
    for j in range(nblk):
        step(j, margin)
        torch.cuda.synchronize()
        want = digest_of(j)
        if want is not None:
            verified.append(sha3(out[: m // 8].cpu().numpy()) == want)
    max_round_dist = float(margin.item())

    # warm-up right before the timed region, so that it starts from a busy GPU rather than from
    # the idle gap of the host-side check above
    for i in range(warmup):
        step(i)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for i in range(steps):
        step(i)
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [evs[i].elapsed_time(evs[i + 1]) for i in range(steps)]
    total_ms = evs[0].elapsed_time(evs[-1])
    want = digest_of((steps - 1) % nblk)
    if want is not None:
        verified.append(sha3(out[: m // 8].cpu().numpy()) == want)
    res = {"value": round(n * steps / (total_ms * 1e-3) / 1e9, 3), "ms_per_step": round(total_ms / steps, 4),
           "median_block_ms": round(float(np.median(per)), 4), "wall_s": round(wall, 3),
           "verified_sha3": (all(verified) if verified else None), "checked_outputs": len(verified),
           "max_round_dist": max_round_dist}
    if with_kernels:
        # per-kernel durations: the engine records HIP events around every launch on the launch
        # stream; this pass repeats the timed region's K blocks with that instrumentation on
        eng.stage_timing(True)
        ia, ib = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ia.record(stream)
        for i in range(steps):
            step(i)
        ib.record(stream)
        ib.synchronize()
        st = eng.stage_times()
        eng.stage_timing(False)
        res["instrumented_ms_per_step"] = round(ia.elapsed_time(ib) / steps, 4)
        kb = kernel_bytes(n, m)
        kernels = []
        for name, (ms, cnt) in st.items():
            avg = ms / cnt
            b = kb.get(name)
            kernels.append({"name": name, "avg_ms": round(avg, 5), "launches": cnt,
                            "alg_bytes": b, "GBs": round(b / avg / 1e6, 1) if b else None})
        kernels.sort(key=lambda k: -k["avg_ms"])
        res["kernels"] = kernels
    if sid:
        eng.seed_drop(sid)
    del keys, seeds, out
    return res


# ------------------------------------------------------------------ batched ----
def run_batched(ctx, steps, warmup, man):
    """The 256-block mix (privacyamplification_amd/workloads.py batched_mix) on device: pa_batch_device
    with a captured graph (the first call tunes new shapes), beside one launch sequence per block
    (pa_block_device) on the same blocks.  Every batched output is checked against its digest."""
    import numpy as np
    from privacyamplification_amd import workloads as W
    torch, P, eng, stream, sptr = ctx.torch, ctx.P, ctx.eng, ctx.stream, ctx.sptr
    spec = W.batched_mix()
    keys, seeds, outa, outb = [], [], [], []
    arr = (P.PaBlock * len(spec))()
    for i, (n, m, ks, ss) in enumerate(spec):
        keys.append(ctx.to_dev(P.synth_bits(ks, n + 1), 1))
        seeds.append(ctx.to_dev(P.synth_bits(ss, n)))
        outa.append(torch.zeros(m // 8 + 16, dtype=torch.uint8, device=ctx.dev))
        outb.append(torch.zeros(m // 8 + 16, dtype=torch.uint8, device=ctx.dev))
        arr[i].key, arr[i].seed = keys[i].data_ptr(), seeds[i].data_ptr()
        arr[i].n, arr[i].m, arr[i].flags, arr[i].out = n, m, P.DEFAULT_FLAGS, outb[i].data_ptr()
    total_bits = sum(n for n, _, _, _ in spec)
    padded_bits = sum(P.block_extents(n, m)[0] for n, m, _, _ in spec)

    def one_per_block():
        for i, (n, m, _, _) in enumerate(spec):
            eng.block_device(keys[i].data_ptr(), seeds[i].data_ptr(), n, m, P.DEFAULT_FLAGS,
                             outa[i].data_ptr(), sptr)

    def batched():
        eng.batch_device(arr, P.BATCH_GRAPH, sptr)

    def timed(fn, k):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(k):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / k

    for _ in range(max(1, warmup)):
        one_per_block()
    eng.batch_device(arr, P.BATCH_GRAPH | P.BATCH_TUNE, sptr)  # tunes new shapes, captures the graph
    for _ in range(max(1, warmup)):
        batched()
    torch.cuda.synchronize()
    ms_single = timed(one_per_block, steps)
    ms_batch = timed(batched, steps)
    # off the clock: every batched output against the oracle digest, the single path against it
    bad_batch = sum(sha3(outb[i][: m // 8].cpu().numpy()) != man[f"c4_{i}"]["sha3"]
                    for i, (n, m, _, _) in enumerate(spec))
    bad_single = sum(not torch.equal(outa[i][: m // 8], outb[i][: m // 8]) for i, (n, m, _, _) in enumerate(spec))
    res = {"value": round(total_bits / (ms_batch * 1e-3) / 1e9, 3), "unit": "Gbit/s",
           "ms_per_step": round(ms_batch, 4), "blocks": len(spec), "bits_per_step": total_bits,
           "transform_bits_per_step": padded_bits,
           "one_launch_per_block_gbit_s": round(total_bits / (ms_single * 1e-3) / 1e9, 3),
           "one_launch_per_block_ms": round(ms_single, 4),
           "speedup_vs_one_launch_per_block": round(ms_single / ms_batch, 3),
           "mismatches_vs_oracle": int(bad_batch), "single_path_mismatches": int(bad_single)}
    del keys, seeds, outa, outb
    return res


# ------------------------------------------------------------------ serving ----
def serving_fixture(ctx, man):
    """The serving set in pinned host memory, resident seeds for the 4 streams, and each distinct
    block's expected output: the engine's output alone, checked against the oracle digest (so
    comparing a served output with it is comparing with the oracle)."""
    import numpy as np
    from privacyamplification_amd import workloads as W
    P, eng = ctx.P, ctx.eng
    streams, dynamic = W.serving_set()
    st, dy, bad = [], [], 0
    for k, (n, m, ts, kss) in enumerate(streams):
        sid = eng.seed_put(P.synth_bits(ts, n), n, m)
        keys, refs = [], []
        for j, ks in enumerate(kss):
            key = eng.pinned_copy(P.synth_bits(ks, n + 1))
            ref = eng.amplify_seeded(key, sid, n, m)
            bad += sha3(ref) != man[f"c5_s{k}_{j}"]["sha3"]
            keys.append(key)
            refs.append(ref)
        st.append((n, m, sid, keys, refs))
    for j, (n, m, ks, ts) in enumerate(dynamic):
        key, seed = eng.pinned_copy(P.synth_bits(ks, n + 1)), eng.pinned_copy(P.synth_bits(ts, n))
        ref = eng.amplify(key, seed, n, m)
        bad += sha3(ref) != man[f"c5_d{j}"]["sha3"]
        dy.append((n, m, key, seed, ref))
    return st, dy, W.serving_mean_n(streams, dynamic), int(bad)


def serve_rate(ctx, st, dy, mean_n, rate, nblk):
    """One Poisson stream (rate in Gbit/s of offered key; inf = every block queued at t = 0)."""
    import queue
    import threading
    import numpy as np
    from privacyamplification_amd import workloads as W
    eng = ctx.eng
    times, kinds, picks = W.serving_arrivals(rate, nblk, mean_n)

    def pick(i):  # (key, n, m, seed or None, resident seed id, expected output)
        p = W.serving_pick(st, dy, kinds[i], picks[i])
        if p[0] == "static":
            n, m, sid, keys, refs = st[p[1]]
            return keys[p[2]], n, m, None, sid, refs[p[2]]
        n, m, key, seed, ref = dy[p[1]]
        return key, n, m, seed, 0, ref

    lats, bad, bits_done, t_end = [], [0], [0], [0.0]
    if math.isinf(rate):
        # saturation: every block is ready at t = 0 (a full input cache); outputs in one pinned
        # arena, blocks go in 256 per pa_serve_submit_many call, outputs checked off the clock
        sel = [pick(i) for i in range(nblk)]
        offs = np.cumsum([0] + [(x[2] // 8 + 255) // 256 * 256 for x in sel])
        arena = eng.host_empty(int(offs[-1]), np.uint8)
        outs = [arena[offs[i]: offs[i] + sel[i][2] // 8] for i in range(nblk)]
        t0 = time.perf_counter()
        tickets = []
        for c in range(0, nblk, 256):
            tickets += eng.serve_submit_many([(k, n, m, outs[c + j], sd, sid)
                                              for j, (k, n, m, sd, sid, _) in enumerate(sel[c:c + 256])])
        for t in tickets:
            lats.append(eng.serve_wait(t, 120 * 10**9))
        t_end[0] = time.perf_counter()
        bits_done[0] = sum(x[1] for x in sel)
        bad[0] = sum(not np.array_equal(o, x[5]) for o, x in zip(outs, sel))
        del outs, arena
    else:
        max_out = max([s_[1] for s_ in st] + [d[1] for d in dy]) // 8 + 8
        free = queue.Queue()
        for _ in range(1024):
            free.put(eng.host_empty(max_out, np.uint8))
        work = queue.Queue()

        def waiter():
            while True:
                item = work.get()
                if item is None:
                    return
                t, buf, ref, n = item
                lats.append(eng.serve_wait(t, 120 * 10**9))
                if not np.array_equal(buf[: ref.size], ref):
                    bad[0] += 1
                bits_done[0] += n
                t_end[0] = time.perf_counter()
                free.put(buf)

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for i in range(nblk):
            while time.perf_counter() - t0 < times[i]:
                pass
            buf = free.get()
            key, n, m, seed, sid, ref = pick(i)
            t = eng.serve_submit(key, n, m, buf, seed=seed, sid=sid)
            work.put((t, buf, ref, n))
        work.put(None)
        th.join()
    span = t_end[0] - t0
    lat_ms = np.array(lats) / 1e6
    return {"offered_gbit_s": "max" if math.isinf(rate) else rate, "blocks": nblk,
            "sustained_gbit_s": round(bits_done[0] / span / 1e9, 3), "span_s": round(span, 3),
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
            "p999_ms": round(float(np.percentile(lat_ms, 99.9)), 3),
            "mismatches_vs_oracle": int(bad[0])}


def plain_capacity(ctx, st, dy, nblk):
    """The plain engine on the serving mix: one synchronous pa_run_block / pa_run_block_seeded call
    per block, the same pinned host buffers, nothing batched or overlapped."""
    import numpy as np
    eng = ctx.eng
    prng = np.random.default_rng(0x9A1)
    bits, t0 = 0, time.perf_counter()
    for _ in range(nblk):
        if prng.random() < 0.75:
            n, m, sid, keys, _ = st[int(prng.integers(0, len(st)))]
            eng.amplify_seeded(keys[int(prng.integers(0, len(keys)))], sid, n, m)
        else:
            n, m, key, seed, _ = dy[int(prng.integers(0, len(dy)))]
            eng.amplify(key, seed, n, m)
        bits += n
    return bits / (time.perf_counter() - t0) / 1e9


def run_serving(ctx, rates, nblk, man, capacity=False):
    st, dy, mean_n, bad_refs = serving_fixture(ctx, man)
    ctx.eng.serve_start(slots=3, max_batch=256, hbm_budget_bytes=64 * 10**9)
    try:
        # warm-up, not counted: the first 500 blocks of the same stream (slot arenas and staging
        # reach their steady-state sizes; the first blocks of a fresh server otherwise wait for them)
        finite = [r for r in rates if not math.isinf(r)]
        serve_rate(ctx, st, dy, mean_n, finite[0] if finite else 32.0, 500)
        results = []
        for r in rates:
            results.append(serve_rate(ctx, st, dy, mean_n, r, nblk))
            progress(serving=results[-1])
    finally:
        ctx.eng.serve_stop()
    out = {"rates": results, "mean_n": mean_n, "reference_outputs_vs_oracle_mismatches": bad_refs,
           "config": "Poisson arrivals, lengths from the batched mix, 4 resident seed streams (75 %) + 25 % "
                     "dynamic seeds, XOR + compress, 64 GB budget, 3 slots, pinned host buffers in and out; "
                     "latency = submission to output written; after a 500-block warm-up of the same stream"}
    if capacity:
        plain = plain_capacity(ctx, st, dy, min(nblk, 4000))
        sat = [r for r in results if r["offered_gbit_s"] == "max"]
        out["capacity"] = {"saturated_gbit_s": sat[0]["sustained_gbit_s"] if sat else None,
                           "plain_one_block_per_call_gbit_s": round(plain, 3),
                           "ratio": round(sat[0]["sustained_gbit_s"] / plain, 3) if sat else None}
    for _, _, sid, _, _ in st:
        ctx.eng.seed_drop(sid)
    return out


# ------------------------------------------------------------------ precision ----
def bench_precision(args):
    """Precision sweep: for n in {2^24, 2^26, 2^28, 2^29}, key Bernoulli(0.5) and (0.9), seed
    Bernoulli(0.5), m = 3n/8: worst |v - rint(v)| before rounding and SHA3 equality with the exact
    digest, for the full-length FP32 transform and for FP32 sub-blocks b in {2^20, 2^22, 2^24}
    (b < n); the FP64 and reference-FP32 margins come from the CPU fixture precision_cpu.json."""
    import numpy as np
    ctx = Ctx(max_n_log2=29)
    torch, P, eng, stream, sptr = ctx.torch, ctx.P, ctx.eng, ctx.stream, ctx.sptr
    man = [b for b in manifest().values() if b["name"].startswith("c3_")]
    cpu_file = ROOT / "tests/golden/precision_cpu.json"
    cpu = json.loads(cpu_file.read_text())["blocks"] if cpu_file.exists() else {}
    rows = []
    for b in man:
        n, m = b["n"], b["m"]
        dk = ctx.to_dev(P.synth_bits(b["key_seed"], n + 1, b["p_key"]), 1)
        ds = ctx.to_dev(P.synth_bits(b["toeplitz_seed"], n))
        out = torch.zeros(m // 8 + 16, dtype=torch.uint8, device=ctx.dev)
        mg = torch.zeros(1, dtype=torch.float32, device=ctx.dev)

        def run(fn):
            out.zero_()
            mg.zero_()
            torch.cuda.synchronize()
            a, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            e.record(stream)
            torch.cuda.synchronize()
            ok = sha3(out[: m // 8].cpu().numpy()) == b["sha3"]
            return {"margin": round(float(mg.item()), 6), "exact": ok, "ms": round(a.elapsed_time(e), 3)}

        row = {"name": b["name"], "n": n, "p_key": b["p_key"]}
        row["fp32_full"] = run(lambda: eng.block_device(dk.data_ptr(), ds.data_ptr(), n, m, P.DEFAULT_FLAGS,
                                                        out.data_ptr(), sptr, mg.data_ptr()))
        for sub in (20, 22, 24):
            if (1 << sub) >= n:
                continue
            row[f"fp32_sub_2^{sub}"] = run(lambda: eng.block_device_split(
                dk.data_ptr(), ds.data_ptr(), n, m, P.DEFAULT_FLAGS, sub, out.data_ptr(), sptr, mg.data_ptr()))
        c = cpu.get(b["name"], {})
        for k in ("fp64_margin", "fp64_exact", "ref_fp32_margin", "ref_fp32_exact"):
            row[k] = c.get(k)
        rows.append(row)
        progress(precision=row)
        del dk, ds, out
    worst = max(r["fp32_full"]["margin"] for r in rows)
    print(json.dumps({
        "metric": "precision sweep: worst |v - rint(v)| before rounding, FP32 full length (n = 2^24..2^29)",
        "value": worst, "unit": "distance to nearest integer (fails at 0.5)", "n_gpus": 1, "steps": len(rows),
        "warmup": 0, "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "n in {2^24, 2^26, 2^28, 2^29}, key p in {0.5, 0.9}, seed p 0.5, m = 3n/8, "
                               "XOR + compress; FP32 full length and FP32 sub-blocks b in {2^20, 2^22, 2^24}"},
        "all_exact": all(v["exact"] for r in rows for k, v in r.items() if k.startswith("fp32_")),
        "rows": rows}), flush=True)
    ctx.eng.close()


# ------------------------------------------------------------------ main ----
def base_line(metric, value, unit, args, ms_per_step, config):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": config}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--mode", choices=("dynamic", "static"), default="dynamic")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("all", "block", "batched", "serving", "precision"), default="all",
                    help="all: the headline block plus static, batched and serving objects (default); block: "
                         "the single block only (--mode, --log2n); batched: 256 uneven blocks per step; serving: "
                         "Poisson arrivals at --rates; precision: rounding margins at n = 2^24..2^29")
    ap.add_argument("--serve-blocks", type=int, default=8000, help="blocks per offered rate (serving)")
    ap.add_argument("--rates", default="2,8,32,max",
                    help="offered loads in Gbit/s for --workload serving; 'max' queues every block at once")
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 512 MiB streaming kernel (reference point for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (blocks are independent: "
              "N GPUs would run N independent engines with no communication)", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "no GPU visible"
    if args.workload == "precision":
        return bench_precision(args)
    man = manifest()
    ctx = Ctx(max_n_log2=max(27, args.log2n))
    if args.copy_calibration:
        # a plain streaming kernel of known traffic: reads 512 MiB, writes 512 MiB per launch
        src = torch.empty(1 << 27, dtype=torch.float32, device=ctx.dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(2):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        del src, dst

    if args.workload == "batched":
        b = run_batched(ctx, args.batched_steps, args.warmup, man)
        line = base_line("batched uneven blocks Gbit/s (256 blocks, n log-uniform 2^16..2^25), bit-exact",
                         b["value"], "Gbit/s", args, b["ms_per_step"],
                         {"workload": "256 uneven blocks per step, n = 32*floor(2^U/32) with U ~ U[16,25], "
                                      "m/n in {1/8,1/4,3/8,1/2}, dynamic seeds, XOR + compress"})
        line["steps"] = args.batched_steps
        line.update({k: v for k, v in b.items() if k not in ("value", "ms_per_step")})
        print(json.dumps(line), flush=True)
        ctx.eng.close()
        return 0 if b["mismatches_vs_oracle"] == 0 else 1
    if args.workload == "serving":
        rates = [math.inf if r.strip() == "max" else float(r) for r in args.rates.split(",")]
        s = run_serving(ctx, rates, args.serve_blocks, man, capacity=any(math.isinf(r) for r in rates))
        finite = [r for r in s["rates"] if r["offered_gbit_s"] != "max"]
        top = finite[-1] if finite else s["rates"][-1]
        line = base_line("served Gbit/s with p50/p99/p99.9 block latency (Poisson arrivals, batched length mix), "
                         "bit-exact", top["sustained_gbit_s"], "Gbit/s", args, None,
                         {"workload": s["config"], "mean_n": s["mean_n"]})
        line["steps"], line["warmup"] = len(rates), 0
        line.update(s)
        print(json.dumps(line), flush=True)
        ctx.eng.close()
        ok = s["reference_outputs_vs_oracle_mismatches"] == 0 and all(r["mismatches_vs_oracle"] == 0
                                                                       for r in s["rates"])
        return 0 if ok else 1

    n = 1 << args.log2n
    m = 3 * n // 8
    mode = args.mode if args.workload == "block" else "dynamic"
    h = run_single(ctx, n, m, mode, args.steps, args.warmup, man)
    progress(headline={k: h[k] for k in ("value", "median_block_ms", "verified_sha3")})
    kernels = h.pop("kernels")
    dominant = kernels[0] if kernels else None
    bytes_block = algorithmic_bytes(n, m, mode)
    block_gbs = bytes_block / (h["median_block_ms"] * 1e-3) / 1e9
    dom_gbs = dominant["GBs"] if dominant else None
    ok = h["verified_sha3"] is not False
    line = base_line("amplified key Gbit/s at 2^27-bit blocks, bit-exact; p99 block latency in serving",
                     h["value"] if ok else None, "Gbit/s", args, h["ms_per_step"],
                     {"workload": f"single block n=2^{args.log2n} bits, m=3n/8, {mode} seed, "
                                  f"XOR key rest + compress, Bernoulli(0.5) key and seed",
                      "n": n, "m": m, "seed_mode": mode, "blocks_per_step": 1})
    line["roofline"] = {
        "bound": "hbm", "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(dom_gbs / HBM_PEAK_GBS, 4) if dom_gbs else None,
        "traffic": read_traffic(dominant["name"], args.log2n, mode) if dominant else None,
        "kernel": dominant["name"] if dominant else None,
        "kernel_alg_bytes": dominant["alg_bytes"] if dominant else None,
        "kernel_avg_ms": dominant["avg_ms"] if dominant else None,
        "block_achieved": round(block_gbs, 1), "block_frac": round(block_gbs / HBM_PEAK_GBS, 4),
        "block_alg_bytes": bytes_block, "median_block_ms": h["median_block_ms"],
        "instrumented_ms_per_step": h["instrumented_ms_per_step"]}
    line["kernels"] = kernels
    line["dominant_kernel"] = dominant["name"] if dominant else None
    line["output_rate_gbit_s"] = round(m * args.steps / (h["ms_per_step"] * args.steps * 1e-3) / 1e9, 3)
    for k in ("verified_sha3", "checked_outputs", "max_round_dist", "wall_s"):
        line[k] = h[k]

    if args.workload == "all":
        s = run_single(ctx, n, m, "static", args.steps, args.warmup, man, with_kernels=False)
        sb = algorithmic_bytes(n, m, "static")
        line["static"] = {"value": s["value"], "unit": "Gbit/s", "ms_per_step": s["ms_per_step"],
                          "median_block_ms": s["median_block_ms"],
                          "block_frac": round(sb / (s["median_block_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "verified_sha3": s["verified_sha3"], "checked_outputs": s["checked_outputs"],
                          "max_round_dist": s["max_round_dist"],
                          "config": "n=2^27, m=3n/8, resident seed 0x5EED, keys 0x1000+i"}
        ok = ok and s["verified_sha3"] is not False
        progress(static=line["static"])
        b = run_batched(ctx, args.batched_steps, min(args.warmup, 3), man)
        b["config"] = ("256 uneven blocks per step, n = 32*floor(2^U/32), U ~ U[16,25], m/n in {1/8,1/4,3/8,1/2}, "
                       "dynamic seeds, XOR + compress; pa_batch_device with a captured graph vs one "
                       "pa_block_device call per block")
        line["batched"] = b
        ok = ok and b["mismatches_vs_oracle"] == 0
        progress(batched=b)
        sv = run_serving(ctx, [32.0], args.serve_blocks, man)
        r = sv["rates"][0]
        line["serving"] = dict(r, unit="Gbit/s", reference_outputs_vs_oracle_mismatches=
                               sv["reference_outputs_vs_oracle_mismatches"], config=sv["config"],
                               mean_n=sv["mean_n"])
        ok = ok and r["mismatches_vs_oracle"] == 0 and sv["reference_outputs_vs_oracle_mismatches"] == 0
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(n, m, legs=args.workload == "all")
    if not ok:
        line["value"] = None
        line["error"] = "an output differs from the oracle digest"
    print(json.dumps(line), flush=True)
    ctx.eng.close()
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())

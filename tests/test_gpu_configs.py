"""The measurement plan's multi-block configurations at their stated workloads, every output against
the committed oracle digests (tests/golden/manifest.json, written by tests/golden/make_manifest.py
from the exact NTT oracle):

  C4  256 uneven blocks in one pa_batch_device call (n log-uniform 2^16..2^25, m/n in {1/8, 1/4,
      3/8, 1/2}), captured graph with the tuner, then replayed (the batched bank of
      PrivacyAmplification_subblocks.cu:1058-1081, 1254-1403, here with uneven blocks);
  C5  a Poisson serving stream at 32 Gbit/s offered: 4 resident seed streams, 25 % dynamic seeds,
      64 GB budget (the receive -> mainloop -> send pipeline, PrivacyAmplification.cu:1686-2103);
  C3  the sub-block (tiled) evaluation at b = 2^20, 2^22, 2^24 of 2^24-, 2^26- and 2^28-bit blocks
      (the LargeBlocksizeExample tiling, LBS/main.cpp:313-523).
Runs on an MI355X (marker gpu)."""
import hashlib
import json
import pathlib

import numpy as np
import pytest

import privacyamplification_amd as P
from privacyamplification_amd import workloads as W

pytestmark = pytest.mark.gpu
ROOT = pathlib.Path(__file__).resolve().parents[1]
MAN = {b["name"]: b for b in json.loads((ROOT / "tests/golden/manifest.json").read_text())["blocks"]}


def say(*a):
    print("[case]", *a, flush=True)


def sha3(b):
    return hashlib.sha3_256(bytes(b)).hexdigest()


def test_c4_256_uneven_blocks_one_batch(engine):
    import torch
    dev = torch.device("cuda:0")
    spec = W.batched_mix()
    assert len(spec) == 256
    say("C4: 256 blocks, n 2^16..2^25, one pa_batch_device call, graph + tune, then a replay")
    keys, seeds, outs = [], [], []
    arr = (P.PaBlock * len(spec))()
    for i, (n, m, ks, ss) in enumerate(spec):
        assert MAN[f"c4_{i}"]["n"] == n and MAN[f"c4_{i}"]["m"] == m
        k = np.concatenate([P.synth_bits(ks, n + 1), np.zeros(1, np.uint32)])
        keys.append(torch.from_numpy(k.view(np.int32)).to(dev))
        seeds.append(torch.from_numpy(P.synth_bits(ss, n).view(np.int32)).to(dev))
        outs.append(torch.zeros(m // 8 + 16, dtype=torch.uint8, device=dev))
        arr[i].key, arr[i].seed = keys[i].data_ptr(), seeds[i].data_ptr()
        arr[i].n, arr[i].m, arr[i].flags, arr[i].out = n, m, P.DEFAULT_FLAGS, outs[i].data_ptr()
    for rep in range(2):
        for o in outs:
            o.zero_()
        engine.batch_device(arr, P.BATCH_GRAPH | P.BATCH_TUNE)
        torch.cuda.synchronize()
        bad = [i for i, (n, m, _, _) in enumerate(spec) if sha3(outs[i][: m // 8].cpu().numpy()) != MAN[f"c4_{i}"]["sha3"]]
        assert not bad, (rep, bad[:10], [spec[i][:2] for i in bad[:10]])


def test_c5_poisson_stream_32gbps(engine):
    """About 1500 blocks arriving by a Poisson process at 32 Gbit/s of offered key, through the
    serving scheduler; every output equals its oracle digest."""
    import time
    streams, dynamic = W.serving_set()
    say("C5: Poisson 32 Gbit/s, 1500 blocks, 4 resident seeds + 25 % dynamic, 64 GB budget")
    st = []
    for k, (n, m, ts, kss) in enumerate(streams):
        sid = engine.seed_put(P.synth_bits(ts, n), n, m)
        st.append((n, m, sid, [engine.pinned_copy(P.synth_bits(ks, n + 1)) for ks in kss]))
    dy = [(n, m, engine.pinned_copy(P.synth_bits(ks, n + 1)), engine.pinned_copy(P.synth_bits(ts, n)))
          for n, m, ks, ts in dynamic]
    mean_n = W.serving_mean_n(streams, dynamic)
    nblk = 1500
    times, kinds, picks = W.serving_arrivals(32.0, nblk, mean_n, master=0xC5)
    engine.serve_start(slots=3, max_batch=256, hbm_budget_bytes=64 * 10**9)
    jobs = []
    try:
        t0 = time.perf_counter()
        for i in range(nblk):
            while time.perf_counter() - t0 < times[i]:
                pass
            p = W.serving_pick(streams, dynamic, kinds[i], picks[i])
            if p[0] == "static":
                n, m, sid, keys = st[p[1]]
                out = np.zeros(m // 8 + 8, np.uint8)
                t = engine.serve_submit(keys[p[2]], n, m, out, sid=sid)
                name = f"c5_s{p[1]}_{p[2]}"
            else:
                n, m, key, seed = dy[p[1]]
                out = np.zeros(m // 8 + 8, np.uint8)
                t = engine.serve_submit(key, n, m, out, seed=seed)
                name = f"c5_d{p[1]}"
            jobs.append((t, out, m, name))
        lats = [engine.serve_wait(t, 120 * 10**9) for t, _, _, _ in jobs]
    finally:
        engine.serve_stop()
        for _, _, sid, _ in st:
            engine.seed_drop(sid)
    bad = [name for _, out, m, name in jobs if sha3(out[: m // 8]) != MAN[name]["sha3"]]
    assert not bad, (len(bad), bad[:10])
    assert len({name for _, _, _, name in jobs}) > 40  # the stream drew most of the set
    print("C5 p50 %.3f ms p99 %.3f ms" % (np.percentile(np.array(lats) / 1e6, 50), np.percentile(np.array(lats) / 1e6, 99)))


C3 = [("c3_n24_p0.5", 20), ("c3_n24_p0.9", 22), ("c3_n26_p0.5", 20), ("c3_n26_p0.9", 22), ("c3_n26_p0.5", 24),
      ("c3_n28_p0.5", 20), ("c3_n28_p0.9", 22), ("c3_n28_p0.5", 24)]


@pytest.mark.parametrize("name,sub", C3)
def test_c3_sub_blocks(engine, name, sub):
    """pa_block_device_split on the precision-sweep blocks: tiles of b = 2^sub bits through the
    batched engine, XOR-accumulated per band; the output equals the full block's oracle digest
    and the FP32 margin stays far from 1/2."""
    import torch
    b = MAN[name]
    n, m = b["n"], b["m"]
    say(f"C3 split {name} b=2^{sub}")
    dev = torch.device("cuda:0")
    key = np.concatenate([P.synth_bits(b["key_seed"], n + 1, b["p_key"]), np.zeros(1, np.uint32)])
    dk = torch.from_numpy(key.view(np.int32)).to(dev)
    ds = torch.from_numpy(P.synth_bits(b["toeplitz_seed"], n).view(np.int32)).to(dev)
    out = torch.zeros(m // 8 + 16, dtype=torch.uint8, device=dev)
    mg = torch.zeros(1, dtype=torch.float32, device=dev)
    engine.block_device_split(dk.data_ptr(), ds.data_ptr(), n, m, P.DEFAULT_FLAGS, sub, out.data_ptr(),
                              None, mg.data_ptr())
    torch.cuda.synchronize()
    assert sha3(out[: m // 8].cpu().numpy()) == b["sha3"]
    assert 0.0 < float(mg.item()) < 0.05

"""Writes manifest.json: SHA3-256 of the oracle's amp_out for the seeded synthetic blocks the GPU
tests and the bench use at sizes too large to recompute inside a test run.

Inputs follow the measurement plan: xoshiro256** seeded per block, Bernoulli(0.5) bits, key of
n+1 bits, seed of n bits, m = 3n/8, XOR and compress on.  Output digests come from the exact NTT
oracle (oracle/pa_oracle.c, cross-checked against the direct GF(2) product at smaller n).
Run: python tests/golden/make_manifest.py   (missing entries only; the 2^28 and 2^29 blocks take
several minutes each on 8 cores)
"""
import hashlib
import json
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402
from privacyamplification_amd import workloads as W  # noqa: E402


def entry(name, n, m, key_seed, toeplitz_seed, p_key=0.5):
    key = O.synth_bits(key_seed, n + 1, p_key)
    seed = O.synth_bits(toeplitz_seed, n)
    t = time.time()
    out = O.amp_out(key, seed, n, m, method=O.NTT)
    print(f"{name}: n=2^{n.bit_length()-1} {time.time()-t:.1f}s", flush=True)
    return {"name": name, "n": n, "m": m, "key_seed": key_seed, "toeplitz_seed": toeplitz_seed,
            "p_key": p_key, "do_xor_key_rest": True, "do_compress": True,
            "sha3": hashlib.sha3_256(out.tobytes()).hexdigest(), "out_bytes": int(out.size)}


def plan():
    """(name, n, m, key_seed, toeplitz_seed, p_key) of every manifest block: the single-block
    configurations, the 256 uneven blocks of the batched workload (c4_*), the blocks a serving
    stream draws from (c5_s<stream>_<key>, c5_d<index>; privacyamplification_amd/workloads.py) and
    the 2^25-bit blocks of the serving DMA test (dma_*)."""
    items = [("pr1_n20", 1 << 20, 3 << 17, 0x11, 0x22, 0.5)]
    n = 1 << 27
    m = 3 * n // 8
    for i in range(2):
        items.append((f"c2_dynamic_{i}", n, m, 0x1000 + i, 0x2000 + i, 0.5))
        items.append((f"c2_static_{i}", n, m, 0x1000 + i, 0x5EED, 0.5))
    # precision sweep sizes; 2^28 and 2^29 exceed the reference's largest block (oracle-only parity)
    for e in (24, 26, 28, 29):
        nn = 1 << e
        for pk in (0.5, 0.9):
            items.append((f"c3_n{e}_p{pk}", nn, 3 * nn // 8, 0x3000 + e, 0x4000 + e, pk))
    for i, (n, m, ks, ts) in enumerate(W.batched_mix()):
        items.append((f"c4_{i}", n, m, ks, ts, 0.5))
    streams, dynamic = W.serving_set()
    for k, (n, m, ts, kss) in enumerate(streams):
        for j, ks in enumerate(kss):
            items.append((f"c5_s{k}_{j}", n, m, ks, ts, 0.5))
    for j, (n, m, ks, ts) in enumerate(dynamic):
        items.append((f"c5_d{j}", n, m, ks, ts, 0.5))
    for i in (2, 3):  # tests/test_gpu_serve.py::test_serve_large_blocks_by_dma
        items.append((f"dma_{i}", 1 << 25, (7 << 22) if i == 2 else (3 << 22), 0xD100 + i, 0xD200 + i, 0.5))
    return items


def main():
    """Computes the entries missing from manifest.json (all of them with --all)."""
    O.build()
    out = pathlib.Path(__file__).with_name("manifest.json")
    have = {}
    if out.exists() and "--all" not in sys.argv:
        have = {b["name"]: b for b in json.loads(out.read_text())["blocks"]}
    items = []
    for name, n, m, ks, ts, pk in plan():
        items.append(have[name] if name in have else entry(name, n, m, ks, ts, pk))
        out.write_text(json.dumps({"generator": "oracle NTT (exact)", "blocks": items}, indent=1) + "\n")
    print("wrote", out)


if __name__ == "__main__":
    main()

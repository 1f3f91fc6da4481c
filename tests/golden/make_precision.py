"""Writes precision_cpu.json: for every precision-sweep block of manifest.json (n = 2^24 .. 2^29,
key Bernoulli(0.5) and (0.9), seed Bernoulli(0.5), m = 3n/8), the CPU side of the sweep:

  fp64_margin      max |v - rint(v)| of the pre-rounding values of an FP64 FFT evaluation
  ref_fp32_margin  the same for the reference's own FP32 pipeline restated in the oracle
                   (binInt2float scaling, DC zeroing, product with pre-multiplication reduction,
                   normalisation and correction; PrivacyAmplification.cu:570-813)
  fp64_exact / ref_fp32_exact   whether those evaluations round to the exact amp_out (SHA3 equal
                   to the NTT digest in manifest.json)

The GPU side (FP32 full length and FP32 sub-blocks) is measured by bench.py --workload precision.
Run: python tests/golden/make_precision.py   (tens of minutes on 8 cores; entries already present
are kept)
"""
import hashlib
import json
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402


def margin(v):
    return float(np.max(np.abs(v - np.rint(v))))


def main():
    O.build()
    man = json.loads((ROOT / "tests/golden/manifest.json").read_text())["blocks"]
    out = pathlib.Path(__file__).with_name("precision_cpu.json")
    have = json.loads(out.read_text())["blocks"] if out.exists() else {}
    for b in man:
        if not b["name"].startswith("c3_") or b["name"] in have:
            continue
        n, m = b["n"], b["m"]
        t = time.time()
        key = O.synth_bits(b["key_seed"], n + 1, b["p_key"])
        seed = O.synth_bits(b["toeplitz_seed"], n)
        ks, _ = O.key_split(key, n, m)
        row = {"n": n, "m": m, "p_key": b["p_key"]}
        for tag, dbl, meth in (("fp64", True, O.FFT64), ("ref_fp32", False, O.FFT32)):
            row[f"{tag}_margin"] = margin(O.fft_values(ks, n - m + 1, seed, n, m, use_double=dbl))
            row[f"{tag}_exact"] = hashlib.sha3_256(O.amp_out(key, seed, n, m, method=meth).tobytes()).hexdigest() == b["sha3"]
        have[b["name"]] = row
        print(b["name"], row, f"{time.time() - t:.0f}s", flush=True)
        out.write_text(json.dumps({"generator": "oracle FFT64 / FFT32 restatement", "blocks": have}, indent=1) + "\n")
    print("wrote", out)


if __name__ == "__main__":
    main()

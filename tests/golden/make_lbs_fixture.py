"""Extract the large-block golden vectors of nicoboss/PrivacyAmplification as a data fixture.

Reads examples/LargeBlocksizeExample/Matrix.h under the reference checkout (given as argv[1],
default /root/reference) as text and writes lbs_matrix.json next to this script:

  key_input, toeplitz_seed_input   16384 bits each (Matrix.h:4-5), stored as hex of the RefBits
                                   words that binTo4byteLittleEndian (main.cpp:220-230) produces
  chunk_sha3                       60 SHA3-256 digests of the per-tile outputs (Matrix.h:7-248)
  intermediate_sha3                70 digests of the running XOR state (Matrix.h:250-531)
  pre_xor_sha3, key_rest_sha3, result_sha3   (Matrix.h:533-552)

Only the numbers are kept; the fixture is data, not a copy of the header.
"""
import json
import pathlib
import re
import sys


def _array_body(text, name):
    m = re.search(re.escape(name) + r"\s*(\[\d*\])+\s*=\s*\{", text)
    if not m:
        raise SystemExit(f"array {name} not found")
    i, depth = m.end() - 1, 0
    for j in range(i, len(text)):
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
            if depth == 0:
                return text[i + 1 : j]
    raise SystemExit(f"unterminated array {name}")


def _ints(body):
    return [int(t, 0) for t in re.findall(r"0x[0-9A-Fa-f]+|\d+", body)]


def _digests(body):
    vals = _ints(body)
    assert len(vals) % 32 == 0
    return [bytes(vals[i : i + 32]).hex() for i in range(0, len(vals), 32)]


def _bits_to_refbits_hex(bits):
    assert len(bits) % 32 == 0 and set(bits) <= {0, 1}
    words = []
    for w in range(len(bits) // 32):
        v = 0
        for b in bits[32 * w : 32 * w + 32]:
            v = (v << 1) | b
        words.append(v)
    # little-endian uint32 memory image, hex
    return b"".join(v.to_bytes(4, "little") for v in words).hex()


def main():
    ref = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    text = (ref / "examples/LargeBlocksizeExample/Matrix.h").read_text()
    key_bits = _ints(_array_body(text, "key_input"))
    seed_bits = _ints(_array_body(text, "toeplitz_seed_input"))
    fixture = {
        "source": "examples/LargeBlocksizeExample/Matrix.h (values only)",
        "sample_size": 16384,
        "chunk_size": 2048,
        "key_input_refbits_le_hex": _bits_to_refbits_hex(key_bits),
        "toeplitz_seed_input_refbits_le_hex": _bits_to_refbits_hex(seed_bits),
        "chunk_sha3": _digests(_array_body(text, "ampOutInData_hash")),
        "intermediate_sha3": _digests(_array_body(text, "amp_out_intermediate_hash")),
        "pre_xor_sha3": _digests(_array_body(text, "amp_out_pre_hash"))[0],
        "key_rest_sha3": _digests(_array_body(text, "key_rest_hash"))[0],
        "result_sha3": _digests(_array_body(text, "amp_out_reult"))[0],
    }
    assert len(key_bits) == 16384 and len(seed_bits) == 16384
    assert len(fixture["chunk_sha3"]) == 60 and len(fixture["intermediate_sha3"]) == 70
    out = pathlib.Path(__file__).with_name("lbs_matrix.json")
    out.write_text(json.dumps(fixture, indent=1) + "\n")
    print(f"wrote {out} ({out.stat().st_size} bytes)")


if __name__ == "__main__":
    main()

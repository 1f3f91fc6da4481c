"""Restatement of the reference's large-block tiling (examples/LargeBlocksizeExample/main.cpp).

TEST INFRASTRUCTURE ONLY.  The example splits a 16384-bit key into 1024x1024 Toeplitz tiles,
sends each tile through the privacy-amplification service at n=2^11 in NoXOR / no-compress mode
and XOR-accumulates the tile outputs per row band; its Matrix.h holds the SHA3 digests of every
tile output and of every running state.  This module reproduces the exact sequence of
(key, seed) tile requests so the fixtures can be checked against any engine.

  tiles()          the request schedule of seedProvider/keyProvider/receiveAmpOut
                   (main.cpp:256-305, 308-368, 377-437): loop 1 over columnNr = 9..0, loop 2 over
                   rowNr = 1..5
  tile_inputs()    GetLocalSeed / GetLocalKey (main.cpp:104-120)
  run_schedule()   XorWithRow accumulation and the final key-rest XOR (main.cpp:122-126, 498-519)
"""
import hashlib

import numpy as np

SAMPLE = 16384          # sample_size = 2^14 (main.cpp:44)
CHUNK = 2048            # chunk_size = 2^11
SIDE = CHUNK // 2       # chunk_side
VERTICAL = SAMPLE // 4 + SAMPLE // 8     # 6144
HORIZONTAL = SAMPLE // 2 + SAMPLE // 8   # 10240
V_CHUNKS = VERTICAL // SIDE              # 6
H_CHUNKS = HORIZONTAL // SIDE            # 10
SIDE_BLOCKS = SIDE // 32                 # 32


def sha3(b):
    return hashlib.sha3_256(bytes(b)).hexdigest()


def words_from_hex(h):
    return np.frombuffer(bytes.fromhex(h), dtype="<u4").astype(np.uint32)


def tiles():
    """Yield (row, key_chunk, seed_word_offset, first_in_column) in the order tiles are sent."""
    r = 0
    for column in range(H_CHUNKS - 1, -1, -1):
        count = min((H_CHUNKS - 1) - column + 1, V_CHUNKS)
        for i, key_nr in enumerate(range(column, column + count)):
            yield i, key_nr, r, i == 0, True
        r += SIDE_BLOCKS
    for row_nr in range(1, V_CHUNKS):
        count = min(HORIZONTAL // SIDE_BLOCKS, V_CHUNKS - row_nr)
        for key_nr in range(count):
            yield row_nr + key_nr, key_nr, r, key_nr == 0, False
        r += SIDE_BLOCKS


def tile_inputs(key_words, seed_words, key_nr, r):
    """GetLocalSeed: seed = S[r+32 .. r+64) || S[r .. r+32) words; GetLocalKey: key = 0 || K-chunk
    (a 1-bit right shift across 33 words)."""
    seed = np.concatenate([seed_words[r + SIDE_BLOCKS: r + 2 * SIDE_BLOCKS],
                           seed_words[r: r + SIDE_BLOCKS]]).astype(np.uint32)
    off = key_words[key_nr * SIDE_BLOCKS: key_nr * SIDE_BLOCKS + SIDE_BLOCKS].astype(np.uint64)
    key = np.zeros(SIDE_BLOCKS + 1, dtype=np.uint64)
    key[0] = off[0] >> 1
    key[1:SIDE_BLOCKS] = ((off[:-1] & 1) << 31) | (off[1:] >> 1)
    key[SIDE_BLOCKS] = (off[SIDE_BLOCKS - 1] & 1) << 31
    return key.astype(np.uint32), seed


def run_schedule(fixture, tile_fn):
    """Drive tile_fn(key33, seed64) -> 128 output bytes through the schedule.

    Returns the list of chunk digests, intermediate digests, the pre-XOR digest, the key-rest
    digest and the final digest, in the order receiveAmpOut checks them."""
    key_words = words_from_hex(fixture["key_input_refbits_le_hex"])
    seed_words = words_from_hex(fixture["toeplitz_seed_input_refbits_le_hex"])
    acc = np.zeros(VERTICAL // 8, dtype=np.uint8)
    chunk_d, inter_d = [], []
    for row, key_nr, r, first, loop1 in tiles():
        if first and loop1:
            inter_d.append(sha3(acc))
        key, seed = tile_inputs(key_words, seed_words, key_nr, r)
        out = np.asarray(tile_fn(key, seed), dtype=np.uint8)
        chunk_d.append(sha3(out))
        acc[row * (SIDE // 8):(row + 1) * (SIDE // 8)] ^= out
        inter_d.append(sha3(acc))
    pre = sha3(acc)
    key_bytes = key_words.astype("<u4").tobytes()
    key_rest = np.frombuffer(key_bytes[HORIZONTAL // 8: HORIZONTAL // 8 + VERTICAL // 8],
                             dtype=np.uint8)
    kr_digest = sha3(key_rest)
    # amp_out_arr[i..i+3] ^= key_rest[i+3..i] : the key rest words byte-swapped (main.cpp:512-518)
    result = acc ^ key_rest.reshape(-1, 4)[:, ::-1].reshape(-1)
    return chunk_d, inter_d, pre, kr_digest, sha3(result), result


def full_block_inputs(fixture):
    """The single XOR-mode block equivalent to the whole tiling: n=16384, m=6144,
    key' = 0 || K (16385 bits), seed' = S rotated left by h=10240 bits."""
    key_words = words_from_hex(fixture["key_input_refbits_le_hex"])
    seed_words = words_from_hex(fixture["toeplitz_seed_input_refbits_le_hex"])
    k = key_words.astype(np.uint64)
    key = np.zeros(k.size + 1, dtype=np.uint64)
    key[0] = k[0] >> 1
    key[1:k.size] = ((k[:-1] & 1) << 31) | (k[1:] >> 1)
    key[k.size] = (k[-1] & 1) << 31
    rot = HORIZONTAL // 32
    seed = np.concatenate([seed_words[rot:], seed_words[:rot]]).astype(np.uint32)
    return key.astype(np.uint32), seed, SAMPLE, VERTICAL

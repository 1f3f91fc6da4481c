"""Sub-block (tiled) evaluation (pa_run_block_split / pa_block_device_split): the Toeplitz product
cut into b x b tiles through the batched engine, XOR-accumulated per band, must equal the oracle and
the full-length engine bit for bit, with a smaller rounding margin.  Runs on an MI355X (gpu)."""
import numpy as np
import pytest

import privacyamplification_amd as P

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("n,m,sub,do_xor", [
    (1 << 12, 3 << 9, 6, True), (1 << 12, 3 << 9, 8, False), (1 << 14, 1 << 12, 10, True),
    (96000, 24000, 10, True), (5 << 12, 1 << 12, 9, True), (1 << 16, 3 << 13, 12, True),
    (1 << 16, 1 << 15, 14, False), (100000 - 100000 % 32, 25024, 11, True)])
def test_split_vs_oracle(engine, oracle, n, m, sub, do_xor):
    say(f"split n={n} m={m} b=2^{sub} xor={do_xor}")
    key = oracle.synth_bits(0xD000 + sub, n + 1)
    if not do_xor:
        key = key[: (n - m) // 32 + 1].copy()
    seed = oracle.synth_bits(0xD100 + sub, n)
    ref = oracle.amp_out(key, seed, n, m, do_xor, True, method=oracle.NTT)
    assert np.array_equal(engine.amplify_split(key, seed, n, m, sub, do_xor), ref)


def test_split_equals_full_length_with_smaller_margin(engine, oracle):
    import torch
    n, m = 1 << 22, 3 << 19
    say("split n=2^22 b=2^16 vs full length, margins")
    dev = torch.device("cuda:0")
    key = oracle.synth_bits(0xE1, n + 1, 0.9)
    seed = oracle.synth_bits(0xE2, n)
    dk = torch.from_numpy(np.concatenate([key, np.zeros(1, np.uint32)]).view(np.int32)).to(dev)
    ds = torch.from_numpy(seed.view(np.int32)).to(dev)
    o_full = torch.zeros(m // 8 + 16, dtype=torch.uint8, device=dev)
    o_split = torch.zeros_like(o_full)
    mg = torch.zeros(2, dtype=torch.float32, device=dev)
    s = torch.cuda.Stream(device=dev)
    engine.block_device(dk.data_ptr(), ds.data_ptr(), n, m, P.DEFAULT_FLAGS, o_full.data_ptr(), s.cuda_stream,
                        mg.data_ptr())
    engine.block_device_split(dk.data_ptr(), ds.data_ptr(), n, m, P.DEFAULT_FLAGS, 16, o_split.data_ptr(),
                              s.cuda_stream, mg.data_ptr() + 4)
    s.synchronize()
    assert torch.equal(o_full, o_split)
    full_margin, split_margin = mg.cpu().tolist()
    assert 0 < split_margin < full_margin < 0.5, (split_margin, full_margin)
    ref = oracle.amp_out(key, seed, n, m, method=oracle.NTT)
    assert np.array_equal(o_split[: m // 8].cpu().numpy(), ref)


def test_split_rejects_bad_shapes(engine, oracle):
    n, m = 1 << 12, 1 << 10
    key, seed = oracle.synth_bits(1, n + 1), oracle.synth_bits(2, n)
    for sub in (5, 12, 13):  # b below 2^6, b >= n
        with pytest.raises(P.EngineError):
            engine.amplify_split(key, seed, n, m, sub)

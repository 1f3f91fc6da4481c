"""pa_block_device_noalloc (allocation-free variant of pa_block_device), called through the C ABI:
bit-identical to pa_block_device on two calls in a row, the second smaller than the first (the
shared scratch keeps the first call's contents), and a block beyond its size limit is refused with
PA_ERR_SIZE and a message naming the limit.  Runs on an MI355X (marker gpu)."""
import ctypes

import numpy as np
import pytest

import privacyamplification_amd as P

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _dev_block(torch, dev, n, m, ks, ss):
    key = np.concatenate([P.synth_bits(ks, n + 1), np.zeros(1, np.uint32)])
    dk = torch.from_numpy(key.view(np.int32)).to(dev)
    ds = torch.from_numpy(P.synth_bits(ss, n).view(np.int32)).to(dev)
    return dk, ds


@pytest.mark.parametrize("first,second", [((1 << 22, 3 << 19, 3), (1 << 20, 3 << 17, 3)),
                                          ((3 << 20, 1 << 20, 2), (96000, 24000, 3)),
                                          ((1 << 21, 7 << 18, 0), (1 << 14, 6144, 1)),
                                          ((1 << 16, 1 << 14, 1), (1 << 10, 1 << 8, 3))])
def test_noalloc_equals_block_device_twice(first, second):
    import torch
    dev = torch.device("cuda:0")
    L = P.load_library()
    eng = P.Engine()
    try:
        for n, m, flags in (first, second):
            say(f"noalloc n={n} m={m} flags={flags}")
            dk, ds = _dev_block(torch, dev, n, m, 0x4A00 + n % 1013, 0x4B00 + n % 1013)
            nb = P.out_bytes(n, m, bool(flags & P.COMPRESS))
            ref = torch.zeros(nb + 64, dtype=torch.uint8, device=dev)
            got = torch.full((nb + 64,), 0x5A, dtype=torch.uint8, device=dev)
            mg = torch.zeros(1, dtype=torch.float32, device=dev)
            eng.block_device(dk.data_ptr(), ds.data_ptr(), n, m, flags, ref.data_ptr())
            rc = L.pa_block_device_noalloc(eng._h, dk.data_ptr(), ds.data_ptr(), n, m, flags, got.data_ptr(), None,
                                           mg.data_ptr())
            assert rc == 0, L.pa_strerror(rc)
            torch.cuda.synchronize()
            assert torch.equal(got[:nb], ref[:nb]), (n, m, flags)
            assert bool((got[nb:] == 0x5A).all())
            assert 0.0 < float(mg.item()) < 0.05
    finally:
        eng.close()


def test_noalloc_refuses_blocks_beyond_its_limit():
    import torch
    dev = torch.device("cuda:0")
    L = P.load_library()
    eng = P.Engine()
    try:
        n, m = 1 << 23, 3 << 20
        say(f"noalloc refuses n={n}")
        dk, ds = _dev_block(torch, dev, n, m, 1, 2)
        out = torch.zeros(m // 8, dtype=torch.uint8, device=dev)
        rc = L.pa_block_device_noalloc(eng._h, dk.data_ptr(), ds.data_ptr(), n, m, P.DEFAULT_FLAGS,
                                       out.data_ptr(), None, None)
        assert rc == -2
        assert b"2^22" in L.pa_strerror(rc)
        torch.cuda.synchronize()
        assert int(out.sum()) == 0  # nothing ran
    finally:
        eng.close()

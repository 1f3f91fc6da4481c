"""The C ABI library loads and exports every entry point include/pa_engine.h declares (CPU only:
no compute calls), and its host-side helpers agree with the oracle."""
import pathlib
import re

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include/pa_engine.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s+(pa_[a-z0-9_]+)\s*\(", text, re.M)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("pa_create", "pa_destroy", "pa_seed_put", "pa_seed_drop", "pa_run_block",
                 "pa_run_block_seeded", "pa_submit", "pa_wait", "pa_block_device",
                 "pa_block_device_seeded", "pa_strerror"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import privacyamplification_amd as P
    L = P.load_library()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_strerror_and_version():
    import privacyamplification_amd as P
    L = P.load_library()
    assert L.pa_strerror(0) == b"ok"
    assert b"size" in L.pa_strerror(-2)
    assert L.pa_version().startswith(b"pa-mi355x")


@pytest.mark.parametrize("p", [0.5, 0.9, 0.1])
def test_product_generator_matches_oracle(oracle, p):
    import privacyamplification_amd as P
    for nbits in (1, 31, 32, 33, 1000, 65537):
        assert np.array_equal(P.synth_bits(0xBEEF + nbits, nbits, p), oracle.synth_bits(0xBEEF + nbits, nbits, p))


def test_workspace_sizes_and_rejections():
    import privacyamplification_amd as P
    n = 1 << 27
    ws = P.workspace_bytes(n, 3 * n // 8)
    assert 1 << 30 < ws < 3 << 30  # two N-complex arrays dominate (2 x 512 MiB)
    assert P.workspace_bytes(1 << 14, 6144) < 1 << 20
    assert P.workspace_bytes(1000, 96) == 0          # not a multiple of 32
    assert P.workspace_bytes(1024, 1024) == 0        # m >= n
    assert P.workspace_bytes(3 << 20, 3 << 18, P.XOR_KEY_REST) == 0  # non-2^k without compression
    assert P.workspace_bytes(1 << 30, 3 << 27) == 0  # beyond the largest served transform (2^29)
    assert 4 << 30 < P.workspace_bytes(1 << 29, 3 << 26) < 6 << 30


def test_no_cpu_fallback_when_library_missing(tmp_path, monkeypatch):
    import privacyamplification_amd.engine as E
    monkeypatch.setattr(E, "_lib", None)
    monkeypatch.setattr(E, "_LIB", tmp_path / "libpa_hip.so")
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        E.load_library()

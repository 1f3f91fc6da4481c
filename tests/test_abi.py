"""The C ABI library loads and exports every entry point include/pa_engine.h declares (CPU only:
no compute calls), its host-side helpers agree with the oracle, and bench.py keeps its contract."""
import pathlib
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_functions(header="pa_engine.h"):
    text = (ROOT / "include" / header).read_text()
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
    # at 2^29: key and seed spectra plus pass B's tiles, 2 GiB each
    assert 6 << 30 < P.workspace_bytes(1 << 29, 3 << 26) < 7 << 30


def test_no_cpu_fallback_when_library_missing(tmp_path, monkeypatch):
    import privacyamplification_amd.engine as E
    monkeypatch.setattr(E, "_lib", None)
    monkeypatch.setattr(E, "_LIB", tmp_path / "libpa_hip.so")
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        E.load_library()


def test_zmtp_library_exports_every_declared_symbol():
    import ctypes
    lib = ROOT / "privacyamplification_amd" / "libpa_zmtp.so"
    if not lib.exists():
        pytest.skip("libpa_zmtp.so not built")
    L = ctypes.CDLL(str(lib))
    names = declared_functions("pa_zmtp.h")
    assert {"pa_zmtp_open", "pa_zmtp_send", "pa_zmtp_recv", "pa_sha3_256"} <= set(names)
    assert not [n for n in names if not hasattr(L, n)]


def _server(tmp_path, cfg_text, *args):
    import subprocess
    exe = ROOT / "privacyamplification_amd" / "pa_server"
    if not exe.exists():
        pytest.skip("pa_server not built")
    cfg = tmp_path / "config.yaml"
    if cfg_text is not None:
        cfg.write_text(cfg_text)
    return subprocess.run([str(exe), "--config", str(cfg), *args], cwd=tmp_path, capture_output=True,
                          text=True, timeout=60)


def test_server_config_errors_exit_like_the_reference(tmp_path):
    """The front-end's exit codes before any device work (PrivacyAmplification.cu:1228-1242, 1415):
    102 for a missing or malformed config.yaml, 400 for an unusable factor_exp."""
    assert _server(tmp_path, None).returncode == 102
    assert _server(tmp_path, "factor_exp 27\n").returncode == 102          # no colon
    assert _server(tmp_path, "do_compress: maybe\n").returncode == 102     # bad boolean
    assert _server(tmp_path, "factor_exp: 5  # too small\n").returncode == 400 & 0xFF  # 8-bit status
    assert _server(tmp_path, "factor_exp: 27\nverify_ampout: true\nverify_ampout_sha3: 'abc'\n").returncode == 102


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it says so and exits non-zero before any GPU work."""
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_bench_algorithmic_bytes_model():
    """The roofline's byte model: 24n + n/4 + m/4 per dynamic block, 20n + n/8 + m/4 static, and the
    row pass moves 12n (two spectra in, R' out)."""
    sys.path.insert(0, str(ROOT))
    import bench
    n, m = 1 << 27, 3 << 24
    assert bench.algorithmic_bytes(n, m, "dynamic") == 24 * n + n / 4 + m / 4
    assert bench.algorithmic_bytes(n, m, "static") == 20 * n + n / 8 + m / 4
    assert bench.kernel_bytes(n, m)["k_rows"] == 12 * n
    assert abs(bench.algorithmic_bytes(n, m, "dynamic") - 3.267e9) < 1e6

"""The ZeroMQ front-end end to end: pa_server runs as a child process; Python peers play the
reference's seed server and key server (PUSH, binding) and amp-out client (PULL, connecting) with
the reference's frame layouts.  Every output must equal the oracle for the seed the reuse rule
selects, in order; ampout.bin must hold the first outputs.  Runs on an MI355X (marker gpu)."""
import pathlib
import struct
import subprocess
import threading

import pytest

from tests import zmtp_peer as Z

pytestmark = pytest.mark.gpu
ROOT = pathlib.Path(__file__).resolve().parents[1]


def test_pa_server_end_to_end(tmp_path, oracle):
    exe = ROOT / "privacyamplification_amd" / "pa_server"
    assert exe.exists(), "pa_server not built"
    e = 16
    n = 1 << e
    ps, pk, po = Z.free_port(), Z.free_port(), Z.free_port()
    (tmp_path / "config.yaml").write_text(f"""
factor_exp: {e}
reuse_seed_amount: 0
vertical_len: {3 * n // 8}
do_xor_key_rest: true
do_compress: true
input_blocks_to_cache: 4
output_blocks_to_cache: 4
show_ampout: -1
show_zeromq_status: false
use_matrix_seed_server: true
address_seed_in: 'tcp://127.0.0.1:{ps}'
use_key_server: true
address_key_in: 'tcp://127.0.0.1:{pk}'
host_ampout_server: true
address_amp_out: 'tcp://127.0.0.1:{po}'
store_first_ampouts_in_file: 2
verify_ampout: false
""")
    # the reuse rule: a seed with reuse k > 0 serves k blocks, 0 one block, -1 every later block
    seeds = [(2, oracle.synth_bits(0xE0, n)), (0, oracle.synth_bits(0xE1, n)), (-1, oracle.synth_bits(0xE2, n))]
    schedule = [0, 0, 1] + [2] * 7
    vbs = [n // 32 * 3 // 8, n // 32 // 4, n // 32 // 2, n // 32 // 8, 700, n // 32 * 3 // 8, 1000, 512, 64, 1500]
    blocks = []
    for i, (si, vb) in enumerate(zip(schedule, vbs)):
        m = 32 * vb
        do_xor = i % 3 != 1
        do_compress = i != 5
        key = oracle.synth_bits(0xF00 + i, n + 1)
        wire = key if do_xor else key[: (n - m) // 32 + 1]  # without the XOR only h+1 key bits travel
        blocks.append((key, seeds[si][1], m, do_xor, do_compress, wire.tobytes()))
    seed_srv = Z.Peer("PUSH", ps, bind=True)
    key_srv = Z.Peer("PUSH", pk, bind=True)
    print("[case] pa_server n=2^16 10 blocks", flush=True)
    proc = subprocess.Popen([str(exe), "--config", "config.yaml", "--blocks", str(len(blocks))], cwd=tmp_path,
                            stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    try:
        seed_srv.accept()
        key_srv.accept()

        def feed():
            for reuse, s in seeds:
                seed_srv.send([struct.pack("<i", reuse), s.tobytes()])
            for _, _, m, do_xor, do_compress, wire in blocks:
                key_srv.send([bytes([do_xor]), bytes([do_compress]), struct.pack("<I", m // 32), wire])

        t = threading.Thread(target=feed, daemon=True)
        t.start()
        out = Z.Peer("PULL", po, bind=False, timeout=60)
        got = [out.recv() for _ in blocks]
        t.join(30)
        rc = proc.wait(timeout=60)
        log = proc.stdout.read()
    finally:
        if proc.poll() is None:
            proc.kill()
        for p in (seed_srv, key_srv):
            p.close()
    assert rc == 0, log
    refs = [oracle.amp_out(k, s, n, m, x, c, method=oracle.NTT) for k, s, m, x, c, _ in blocks]
    for i, (g, r) in enumerate(zip(got, refs)):
        assert len(g) == 1 and g[0] == r.tobytes(), (i, len(g[0]), r.size)
    assert (tmp_path / "ampout.bin").read_bytes() == refs[0].tobytes() + refs[1].tobytes()


def _file_config(tmp_path, e, m, extra):
    (tmp_path / "config.yaml").write_text(f"""
factor_exp: {e}
reuse_seed_amount: -1
vertical_len: {m}
do_xor_key_rest: true
do_compress: true
show_ampout: 0
use_matrix_seed_server: false
toeplitz_seed_path: 'seed.bin'
use_key_server: false
keyfile_path: 'key.bin'
host_ampout_server: false
store_first_ampouts_in_file: -1
{extra}
""")


def test_pa_server_file_inputs_and_verify(tmp_path, oracle):
    """File inputs (keyfile_path, toeplitz_seed_path), ampout.bin for every block
    (store_first_ampouts_in_file -1), and verify_ampout against a SHA3-256 digest: the right digest
    passes every block, a wrong one ends the process with the reference's code 101."""
    import hashlib
    exe = ROOT / "privacyamplification_amd" / "pa_server"
    e, m = 18, 3 << 15
    n = 1 << e
    key = oracle.synth_bits(0xA1, n + 1)
    seed = oracle.synth_bits(0xA2, n)
    (tmp_path / "key.bin").write_bytes(key.tobytes())
    (tmp_path / "seed.bin").write_bytes(seed.tobytes())
    ref = oracle.amp_out(key, seed, n, m, method=oracle.NTT).tobytes()
    digest = hashlib.sha3_256(ref).hexdigest()
    print("[case] pa_server files + verify n=2^18", flush=True)
    _file_config(tmp_path, e, m, f"verify_ampout: true\nverify_ampout_sha3: '{digest}'")
    r = subprocess.run([str(exe), "--blocks", "3"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("VERIFIED!") == 3 and r.stdout.count("Blocktime:") == 3
    assert (tmp_path / "ampout.bin").read_bytes() == ref * 3
    bad = ("0" if digest[0] != "0" else "1") + digest[1:]
    _file_config(tmp_path, e, m, f"verify_ampout: true\nverify_ampout_sha3: '{bad}'")
    r = subprocess.run([str(exe), "--blocks", "3"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 101 and "VERIFICATION FAILED!" in r.stdout, r.stdout + r.stderr


def test_pa_server_speedtest(tmp_path):
    """argv speedtest in the reference's output format: d[i,j,k]=Mbit/s for a fresh seed per block
    (i=0) and a resident seed (i=1), n = 2^j for j = 10..27, 10 timed blocks each."""
    import re
    exe = ROOT / "privacyamplification_amd" / "pa_server"
    (tmp_path / "config.yaml").write_text("show_zeromq_status: false\n")
    print("[case] pa_server speedtest", flush=True)
    r = subprocess.run([str(exe), "speedtest"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    d = {tuple(map(int, k.split(","))): float(v) for k, v in re.findall(r"^d\[([\d,]+)\]=([\d.]+)$", r.stdout, re.M)}
    assert set(d) == {(i, j, k) for i in (0, 1) for j in range(10, 28) for k in range(10)}
    assert all(v > 0 for v in d.values())
    print("speedtest 2^27 Mbit/s: dynamic %.0f, resident %.0f" % (
        sorted(d[0, 27, k] for k in range(10))[5], sorted(d[1, 27, k] for k in range(10))[5]))


def test_pa_server_verify_pool_2_24(tmp_path, oracle):
    """verify_ampout on a pool of verify_ampout_threads workers (PrivacyAmplification.cu:1160-1180):
    12 blocks of 2^24 bits from files, every output hashed off the send path against the right
    digest; every block prints VERIFIED! and ampout.bin holds them all."""
    import hashlib
    exe = ROOT / "privacyamplification_amd" / "pa_server"
    e = 24
    n, m = 1 << e, 3 << 21
    key = oracle.synth_bits(0xA5, n + 1)
    seed = oracle.synth_bits(0xA6, n)
    (tmp_path / "key.bin").write_bytes(key.tobytes())
    (tmp_path / "seed.bin").write_bytes(seed.tobytes())
    ref = oracle.amp_out(key, seed, n, m, method=oracle.NTT).tobytes()
    digest = hashlib.sha3_256(ref).hexdigest()
    print("[case] pa_server verify pool n=2^24 12 blocks 4 threads", flush=True)
    _file_config(tmp_path, e, m, f"verify_ampout: true\nverify_ampout_threads: 4\nverify_ampout_sha3: '{digest}'\n"
                                 "output_blocks_to_cache: 6")
    r = subprocess.run([str(exe), "--blocks", "12"], cwd=tmp_path, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert r.stdout.count("VERIFIED!") == 12 and "FAILED" not in r.stdout
    assert (tmp_path / "ampout.bin").read_bytes() == ref * 12


def test_pa_server_verify_failure_ends_unbounded_run(tmp_path, oracle):
    """Without --blocks the block loop runs until its inputs end; file inputs never end, so only
    the fatal path stops it.  A wrong verify digest must end the process with code 101 (the
    reference exits at once, PrivacyAmplification.cu:1136) instead of waiting for more keys."""
    import hashlib
    exe = ROOT / "privacyamplification_amd" / "pa_server"
    e, m = 16, 3 << 13
    n = 1 << e
    key = oracle.synth_bits(0xA7, n + 1)
    seed = oracle.synth_bits(0xA8, n)
    (tmp_path / "key.bin").write_bytes(key.tobytes())
    (tmp_path / "seed.bin").write_bytes(seed.tobytes())
    digest = hashlib.sha3_256(b"not the output").hexdigest()
    print("[case] pa_server unbounded run, wrong digest -> 101", flush=True)
    _file_config(tmp_path, e, m, f"verify_ampout: true\nverify_ampout_sha3: '{digest}'\n"
                                 "input_blocks_to_cache: 2\noutput_blocks_to_cache: 2\nstore_first_ampouts_in_file: 0")
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 101 and "VERIFICATION FAILED!" in r.stdout, r.stdout[-2000:] + r.stderr

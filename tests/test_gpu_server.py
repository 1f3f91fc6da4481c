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

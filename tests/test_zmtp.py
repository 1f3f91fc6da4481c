"""The ZeroMQ front-end's host code on the CPU: SHA3-256 against hashlib, and the ZMTP/3.0 endpoints
of libpa_zmtp.so against a Python peer written from RFC 23 (greeting and READY bytes on the wire,
multipart messages with short and long frames, both roles, reconnection)."""
import ctypes
import hashlib
import pathlib
import socket
import threading

import numpy as np
import pytest

from tests import zmtp_peer as Z

ROOT = pathlib.Path(__file__).resolve().parents[1]
LIB = ROOT / "privacyamplification_amd" / "libpa_zmtp.so"


@pytest.fixture(scope="module")
def zl():
    if not LIB.exists():
        pytest.skip("libpa_zmtp.so not built")
    L = ctypes.CDLL(str(LIB))
    L.pa_zmtp_open.restype = ctypes.c_void_p
    L.pa_zmtp_open.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    L.pa_zmtp_send.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64),
                               ctypes.c_uint32]
    L.pa_zmtp_recv.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]
    L.pa_zmtp_part.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p),
                               ctypes.POINTER(ctypes.c_uint64)]
    L.pa_zmtp_stop.argtypes = [ctypes.c_void_p]
    L.pa_zmtp_close.argtypes = [ctypes.c_void_p]
    L.pa_sha3_256.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_char_p]
    return L


def _send(zl, h, parts):
    lens = (ctypes.c_uint64 * len(parts))(*[len(p) for p in parts])
    keep = [ctypes.c_char_p(p) for p in parts]
    arr = (ctypes.c_void_p * len(parts))(*[ctypes.cast(k, ctypes.c_void_p).value for k in keep])
    assert zl.pa_zmtp_send(h, arr, lens, len(parts)) == 0


def _recv(zl, h):
    n = ctypes.c_uint32()
    assert zl.pa_zmtp_recv(h, ctypes.byref(n)) == 0
    out = []
    for i in range(n.value):
        p, ln = ctypes.c_void_p(), ctypes.c_uint64()
        assert zl.pa_zmtp_part(h, i, ctypes.byref(p), ctypes.byref(ln)) == 0
        out.append(ctypes.string_at(p.value, ln.value) if ln.value else b"")
    return out


@pytest.mark.parametrize("n", [0, 1, 135, 136, 137, 272, 1000, 1 << 20])
def test_sha3_256_matches_hashlib(zl, n):
    data = np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8).tobytes()
    out = ctypes.create_string_buffer(32)
    zl.pa_sha3_256(data, n, out)
    assert out.raw == hashlib.sha3_256(data).digest()


def test_greeting_and_ready_bytes_on_the_wire(zl):
    """Our connecting PULL endpoint must open with the RFC 23 greeting (client) and a READY command
    naming its socket type."""
    port = Z.free_port()
    ls = socket.socket()
    ls.setsockopt(socket.SOL_SOCKET, socket.SO_REUSEADDR, 1)
    ls.bind(("127.0.0.1", port))
    ls.listen(1)
    h = zl.pa_zmtp_open(b"PULL", f"tcp://127.0.0.1:{port}".encode(), 0)
    got = {}

    def run():
        n = ctypes.c_uint32()
        got["rc"] = zl.pa_zmtp_recv(h, ctypes.byref(n))

    t = threading.Thread(target=run)
    t.start()
    conn, _ = ls.accept()
    conn.settimeout(10)
    g = b""
    while len(g) < 64:
        g += conn.recv(64 - len(g))
    assert g[0] == 0xFF and g[9] == 0x7F and g[10] == 3 and g[12:32] == b"NULL" + b"\x00" * 16 and g[32] == 0
    conn.sendall(Z.greeting(True))
    want = bytes([4, 26]) + b"\x05READY" + b"\x0bSocket-Type" + b"\x00\x00\x00\x04PULL"
    r = b""
    while len(r) < len(want):
        r += conn.recv(len(want) - len(r))
    assert r == want
    conn.sendall(Z.ready("PUSH"))
    conn.sendall(bytes([0, 3]) + b"abc")  # one message so the receive returns
    t.join(10)
    assert got["rc"] == 0
    zl.pa_zmtp_stop(h)
    zl.pa_zmtp_close(h)
    conn.close()
    ls.close()


def test_push_bind_to_python_pull(zl):
    port = Z.free_port()
    h = zl.pa_zmtp_open(b"PUSH", f"tcp://127.0.0.1:{port}".encode(), 1)
    msgs = [[b"x"], [b"\x01", b"\x00", b"\x10\x00\x00\x00", bytes(range(256)) * 5], [b"", b"z" * 70000]]
    t = threading.Thread(target=lambda: [_send(zl, h, m) for m in msgs])
    t.start()
    peer = Z.Peer("PULL", port, bind=False)
    assert peer.peer_ready.endswith(b"PUSH")
    for m in msgs:
        assert peer.recv() == m
    t.join(10)
    peer.close()
    zl.pa_zmtp_stop(h)
    zl.pa_zmtp_close(h)


def test_python_push_bind_to_pull_connect_and_reconnect(zl):
    """The engine's PULL endpoints connect to binding peers (the seed and key servers), on their
    first receive; a peer that restarts is reconnected and its next message still arrives whole."""
    port = Z.free_port()
    h = zl.pa_zmtp_open(b"PULL", f"tcp://127.0.0.1:{port}".encode(), 0)
    seed = np.arange(1 << 14, dtype=np.uint32).tobytes()
    for msg in ([np.int32(-1).tobytes(), seed], [b"\x02" * 4, b"\x03" * 300]):
        got = {}
        t = threading.Thread(target=lambda: got.setdefault("m", _recv(zl, h)))  # (re)connects while waiting
        t.start()
        srv = Z.Peer("PUSH", port, bind=True)
        srv.accept()
        srv.send(msg)
        t.join(20)
        assert got.get("m") == msg
        srv.close()
    zl.pa_zmtp_stop(h)
    zl.pa_zmtp_close(h)


def test_bad_addresses_are_refused(zl):
    for a in (b"ipc://x", b"tcp://127.0.0.1", b"tcp://:5", b"tcp://h:99999", b"tcp://h:12ab"):
        assert not zl.pa_zmtp_open(b"PUSH", a, 1)
    assert not zl.pa_zmtp_open(b"REQ", b"tcp://127.0.0.1:5555", 1)

"""A ZMTP/3.0 (RFC 23, NULL security) PUSH/PULL peer in plain Python sockets: the test-side stand-in
for the reference's libzmq peers (seed server, key server, amp-out client).

Wire format: greeting = %xFF, 8 padding octets, %x7F, version 3.0, mechanism "NULL" padded to 20
octets, as-server octet, 31 filler octets; then a READY command carrying Socket-Type; then frames
flags (MORE 1, LONG 2, COMMAND 4) | size (1 octet, or 8 big-endian octets when LONG) | body.
"""
import socket
import struct
import time

MORE, LONG, COMMAND = 1, 2, 4


def greeting(as_server):
    return (b"\xff" + b"\x00" * 7 + b"\x01" + b"\x7f" + bytes([3, 0]) + b"NULL" + b"\x00" * 16
            + bytes([1 if as_server else 0]) + b"\x00" * 31)


def ready(socket_type):
    body = b"\x05READY" + bytes([11]) + b"Socket-Type" + struct.pack(">I", len(socket_type)) + socket_type.encode()
    return bytes([COMMAND, len(body)]) + body


class Peer:
    """One connected ZMTP peer: bind=True listens on 127.0.0.1:port and accepts one connection."""

    def __init__(self, role, port, bind, timeout=30.0):
        self.role = role
        if bind:
            ls = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
            ls.setsockopt(socket.SOL_SOCKET, socket.SO_REUSEADDR, 1)
            ls.bind(("127.0.0.1", port))
            ls.listen(1)
            ls.settimeout(timeout)
            self.listener = ls
            self.sock = None
        else:
            self.listener = None
            deadline = time.time() + timeout
            while True:
                try:
                    self.sock = socket.create_connection(("127.0.0.1", port), timeout=timeout)
                    break
                except OSError:
                    if time.time() > deadline:
                        raise
                    time.sleep(0.05)
            self._handshake(False)

    def accept(self):
        self.sock, _ = self.listener.accept()
        self.sock.settimeout(self.listener.gettimeout())
        self._handshake(True)
        return self

    def _read(self, n):
        buf = b""
        while len(buf) < n:
            chunk = self.sock.recv(n - len(buf))
            if not chunk:
                raise ConnectionError("peer closed")
            buf += chunk
        return buf

    def _handshake(self, as_server):
        self.sock.sendall(greeting(as_server))
        self.peer_greeting = self._read(64)
        self.sock.sendall(ready(self.role))
        flags, body = self._frame()
        assert flags & COMMAND and body.startswith(b"\x05READY"), (flags, body[:16])
        self.peer_ready = body

    def _frame(self):
        flags = self._read(1)[0]
        size = struct.unpack(">Q", self._read(8))[0] if flags & LONG else self._read(1)[0]
        return flags, self._read(size)

    def send(self, parts):
        out = b""
        for i, p in enumerate(parts):
            more = MORE if i + 1 < len(parts) else 0
            if len(p) > 255:
                out += bytes([more | LONG]) + struct.pack(">Q", len(p)) + p
            else:
                out += bytes([more, len(p)]) + p
        self.sock.sendall(out)

    def recv(self):
        parts = []
        while True:
            flags, body = self._frame()
            if flags & COMMAND:
                continue
            parts.append(body)
            if not flags & MORE:
                return parts

    def close(self):
        for s in (self.sock, self.listener):
            if s is not None:
                s.close()


def free_port():
    s = socket.socket()
    s.bind(("127.0.0.1", 0))
    p = s.getsockname()[1]
    s.close()
    return p

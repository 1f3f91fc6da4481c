"""ctypes front of libpa_hip.so (the C ABI of include/pa_engine.h).

The Engine class mirrors the reference's per-block contract (mainloop's inputs and output,
PrivacyAmplification.cu:1686-2103): a key of n+1 bits (n/32+1 RefBits words; h/32+1 words without
key-rest XOR), an n-bit Toeplitz seed (n/32 words), vertical_len m and the two per-block flags
do_xor_key_rest / do_compress.  The output is the amp_out byte string the reference sends on its
amp-out socket (m/8 bytes, or n/16 bytes without compression).

There is no fallback: if the HIP library is missing or no GPU is present, calls raise.
"""
import ctypes
import os
import pathlib
import subprocess
import weakref

import numpy as np

XOR_KEY_REST = 1
COMPRESS = 2
DEFAULT_FLAGS = XOR_KEY_REST | COMPRESS

_PKG = pathlib.Path(__file__).resolve().parent
_LIB = _PKG / "libpa_hip.so"
_lib = None

_STATUS = {
    0: "ok", -1: "invalid argument", -2: "block size outside the supported family",
    -3: "HIP runtime error", -4: "device memory or HBM budget exhausted", -5: "timed out",
    -6: "unknown seed or ticket", -7: "unsupported request", -8: "self-test failed",
}

SELFTESTS = ("CalculateCorrectionFloat", "SetFirstElementToZero", "ElementWiseProduct", "BinInt2float",
             "ToBinaryArray")


class EngineError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = _lib.pa_strerror(status).decode() if _lib is not None else _STATUS.get(status, "?")
        super().__init__(f"{what}: {msg} (status {status})" if what else msg)


class PaConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("max_n_log2", ctypes.c_uint32),
        ("reduction_exp", ctypes.c_uint32),
        ("pre_mul_reduction_exp", ctypes.c_uint32),
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("out_reverse_endian", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class PaBlock(ctypes.Structure):
    """One block of a batch (pa_block): device pointers for batch_device, host for run_batch."""
    _fields_ = [
        ("key", ctypes.c_void_p),
        ("seed", ctypes.c_void_p),
        ("seed_id", ctypes.c_uint64),
        ("n", ctypes.c_uint32),
        ("m", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("out", ctypes.c_void_p),
        ("margin", ctypes.c_void_p),
    ]


BATCH_GRAPH = 1
BATCH_TUNE = 2


class PaServeConfig(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_uint32),
        ("max_batch", ctypes.c_uint32),
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("batch_flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


def library_path():
    return _LIB


def build(force=False):
    """Compile libpa_hip.so for gfx950 with the Makefile in csrc/ (hipcc cross-compiles; no GPU needed)."""
    cmd = ["make", "-C", str(_PKG / "csrc")]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True)
    return _LIB


def load_library():
    """Load the in-tree HIP library (or the one PA_HIP_LIBRARY names, a build of the same sources
    with other options); raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = pathlib.Path(os.environ.get("PA_HIP_LIBRARY", _LIB))
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it with privacyamplification_amd.build() "
                           "(there is no CPU fallback)")
    # One HIP runtime per process: torch ships its own libamdhip64 (same soname).  Loading torch
    # first makes this library bind to that runtime, so device pointers and stream handles from
    # torch are valid here; loaded the other way round, torch would start a second runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(str(path))
    u32p, u8p, vp = ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint8), ctypes.c_void_p
    eng = ctypes.c_void_p
    L.pa_config_default.argtypes = [ctypes.POINTER(PaConfig)]
    L.pa_create.argtypes = [ctypes.POINTER(PaConfig), ctypes.POINTER(eng)]
    L.pa_destroy.argtypes = [eng]
    L.pa_seed_put.argtypes = [eng, u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32,
                              ctypes.POINTER(ctypes.c_uint64)]
    L.pa_seed_drop.argtypes = [eng, ctypes.c_uint64]
    L.pa_run_block.argtypes = [eng, u32p, u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, u8p]
    L.pa_run_block_seeded.argtypes = [eng, u32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, u8p]
    L.pa_submit.argtypes = [eng, u32p, u32p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                            ctypes.c_uint32, u8p, ctypes.POINTER(ctypes.c_uint64)]
    L.pa_wait.argtypes = [eng, ctypes.c_uint64, ctypes.c_int64]
    L.pa_block_device.argtypes = [eng, vp, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp,
                                  vp, vp]
    L.pa_block_device_seeded.argtypes = [eng, vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32, vp, vp, vp]
    L.pa_block_device_noalloc.argtypes = [eng, vp, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp,
                                          vp, vp]
    L.pa_batch_device.argtypes = [eng, ctypes.POINTER(PaBlock), ctypes.c_uint32, ctypes.c_uint32, vp]
    L.pa_run_batch.argtypes = [eng, ctypes.POINTER(PaBlock), ctypes.c_uint32, ctypes.c_uint32]
    L.pa_block_device_split.argtypes = [eng, vp, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32, vp, vp, vp]
    L.pa_run_block_split.argtypes = [eng, u32p, u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_uint32, u8p]
    L.pa_host_alloc.argtypes = [eng, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]
    L.pa_host_free.argtypes = [eng, ctypes.c_void_p]
    L.pa_serve_start.argtypes = [eng, ctypes.POINTER(PaServeConfig)]
    L.pa_serve_submit.argtypes = [eng, ctypes.POINTER(PaBlock), ctypes.POINTER(ctypes.c_uint64)]
    L.pa_serve_submit_many.argtypes = [eng, ctypes.POINTER(PaBlock), ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_uint64)]
    L.pa_serve_wait.argtypes = [eng, ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64)]
    L.pa_serve_stop.argtypes = [eng]
    L.pa_stage_timing.argtypes = [eng, ctypes.c_int]
    L.pa_stage_times.argtypes = [eng, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    L.pa_workspace_bytes.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
    L.pa_workspace_bytes.restype = ctypes.c_uint64
    L.pa_block_extents.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64),
                                   ctypes.POINTER(ctypes.c_uint64)]
    L.pa_synth_bits.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64, u32p]
    L.pa_selftest.argtypes = [eng, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint32]
    L.pa_strerror.argtypes = [ctypes.c_int]
    L.pa_strerror.restype = ctypes.c_char_p
    L.pa_version.restype = ctypes.c_char_p
    for name in ("pa_config_default", "pa_create", "pa_destroy", "pa_seed_put", "pa_seed_drop",
                 "pa_run_block", "pa_run_block_seeded", "pa_submit", "pa_wait", "pa_block_device",
                 "pa_block_device_seeded", "pa_synth_bits", "pa_stage_timing", "pa_stage_times",
                 "pa_batch_device", "pa_run_batch", "pa_serve_start", "pa_serve_submit",
                 "pa_serve_submit_many",
                 "pa_serve_wait", "pa_serve_stop", "pa_host_alloc", "pa_host_free",
                 "pa_block_device_split", "pa_run_block_split", "pa_block_extents", "pa_selftest",
                 "pa_block_device_noalloc"):
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _check(status, what):
    if status != 0:
        raise EngineError(status, what)


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def synth_bits(seed, nbits, p=0.5):
    """Synthetic RefBits (xoshiro256**, Bernoulli(p)) from the product library's generator."""
    L = load_library()
    out = np.zeros((nbits + 31) // 32, dtype=np.uint32)
    _check(L.pa_synth_bits(seed, p, nbits, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
           "pa_synth_bits")
    return out


def workspace_bytes(n, m, flags=DEFAULT_FLAGS):
    return int(load_library().pa_workspace_bytes(n, m, flags))


def out_bytes(n, m, do_compress=True):
    return m // 8 if do_compress else n // 16


BUFFERS = ("key", "seed", "out", "spec_key", "spec_seed", "bits_key", "bits_seed", "nib", "counts", "tiles")


def block_extents(n, m, flags=DEFAULT_FLAGS):
    """pa_block_extents: (L, touched, alloc) for a block shape, touched / alloc as {buffer: bytes}
    (bytes its kernels index at most / bytes every allocation path gives it).  Host only."""
    L = load_library()
    tl = ctypes.c_uint32()
    t = (ctypes.c_uint64 * len(BUFFERS))()
    a = (ctypes.c_uint64 * len(BUFFERS))()
    _check(L.pa_block_extents(n, m, flags, ctypes.byref(tl), t, a), "pa_block_extents")
    return tl.value, dict(zip(BUFFERS, list(t))), dict(zip(BUFFERS, list(a)))


class Engine:
    """One engine per GPU (pa_create / pa_destroy)."""

    def __init__(self, device=0, max_n_log2=27, hbm_budget_bytes=64 << 30, out_reverse_endian=True):
        L = load_library()
        cfg = PaConfig()
        L.pa_config_default(ctypes.byref(cfg))
        cfg.device, cfg.max_n_log2 = device, max_n_log2
        cfg.hbm_budget_bytes = hbm_budget_bytes
        cfg.out_reverse_endian = 1 if out_reverse_endian else 0
        self._h = ctypes.c_void_p()
        _check(L.pa_create(ctypes.byref(cfg), ctypes.byref(self._h)), "pa_create")
        self._L = L
        self._pinned = []  # finalizers of host_empty arrays (must not outlive the engine)

    def close(self):
        if self._h:
            for f in getattr(self, "_pinned", []):  # pa_destroy frees every pinned buffer itself
                f.detach()
            self._pinned = []
            self._L.pa_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def flags(do_xor_key_rest=True, do_compress=True):
        return (XOR_KEY_REST if do_xor_key_rest else 0) | (COMPRESS if do_compress else 0)

    def amplify(self, key, seed, n, m, do_xor_key_rest=True, do_compress=True):
        """One block, host buffers in and out; returns the amp_out bytes as a uint8 array."""
        out = np.zeros(out_bytes(n, m, do_compress) + 8, dtype=np.uint8)
        k, kp = _u32(key)
        s, sp = _u32(seed)
        _check(self._L.pa_run_block(self._h, kp, sp, n, m, self.flags(do_xor_key_rest, do_compress),
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))), "pa_run_block")
        return out[: out_bytes(n, m, do_compress)]

    def seed_put(self, seed, n, m, reuse=-1):
        s, sp = _u32(seed)
        sid = ctypes.c_uint64()
        _check(self._L.pa_seed_put(self._h, sp, n, m, reuse, ctypes.byref(sid)), "pa_seed_put")
        return sid.value

    def seed_drop(self, sid):
        _check(self._L.pa_seed_drop(self._h, sid), "pa_seed_drop")

    def amplify_seeded(self, key, sid, n, m, do_xor_key_rest=True, do_compress=True):
        out = np.zeros(out_bytes(n, m, do_compress) + 8, dtype=np.uint8)
        k, kp = _u32(key)
        _check(self._L.pa_run_block_seeded(self._h, kp, sid, n, m,
                                           self.flags(do_xor_key_rest, do_compress),
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))),
               "pa_run_block_seeded")
        return out[: out_bytes(n, m, do_compress)]

    def submit(self, key, seed, n, m, do_xor_key_rest=True, do_compress=True, sid=0):
        """Asynchronous block; returns (ticket, out, keepalive).  Call wait(ticket) before reading out."""
        out = np.zeros(out_bytes(n, m, do_compress) + 8, dtype=np.uint8)
        k, kp = _u32(key)
        if seed is not None:
            s, sp = _u32(seed)
        else:
            s, sp = None, None
        t = ctypes.c_uint64()
        _check(self._L.pa_submit(self._h, kp, sp, sid, n, m, self.flags(do_xor_key_rest, do_compress),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), ctypes.byref(t)),
               "pa_submit")
        return t.value, out, (k, s)

    def wait(self, ticket, timeout_ns=-1):
        _check(self._L.pa_wait(self._h, ticket, timeout_ns), "pa_wait")

    def block_device(self, d_key, d_seed, n, m, flags, d_out, stream=None, d_margin=None):
        """Device-resident block: integer device addresses (e.g. torch tensor data_ptr())."""
        _check(self._L.pa_block_device(self._h, d_key, d_seed, n, m, flags, d_out, stream, d_margin),
               "pa_block_device")

    def block_device_noalloc(self, d_key, d_seed, n, m, flags, d_out, stream=None, d_margin=None):
        """pa_block_device without any allocation (scratch in a fixed device array; n <= 2^22)."""
        _check(self._L.pa_block_device_noalloc(self._h, d_key, d_seed, n, m, flags, d_out, stream, d_margin),
               "pa_block_device_noalloc")

    def block_device_seeded(self, d_key, sid, n, m, flags, d_out, stream=None, d_margin=None):
        _check(self._L.pa_block_device_seeded(self._h, d_key, sid, n, m, flags, d_out, stream, d_margin),
               "pa_block_device_seeded")

    def run_batch(self, blocks, batch_flags=0):
        """Host-buffer batch (pa_run_batch).  blocks: dicts with key, seed (or None with sid), n, m,
        and optional do_xor_key_rest / do_compress / sid.  Returns the amp_out arrays in order."""
        arr = (PaBlock * len(blocks))()
        keep, outs = [], []
        for i, b in enumerate(blocks):
            fl = self.flags(b.get("do_xor_key_rest", True), b.get("do_compress", True))
            k, _ = _u32(b["key"])
            keep.append(k)
            arr[i].key = k.ctypes.data
            if b.get("seed") is not None:
                sd, _ = _u32(b["seed"])
                keep.append(sd)
                arr[i].seed = sd.ctypes.data
            arr[i].seed_id = b.get("sid", 0)
            arr[i].n, arr[i].m, arr[i].flags = b["n"], b["m"], fl
            o = np.zeros(out_bytes(b["n"], b["m"], fl & COMPRESS) + 8, dtype=np.uint8)
            outs.append(o)
            arr[i].out = o.ctypes.data
        _check(self._L.pa_run_batch(self._h, arr, len(blocks), batch_flags), "pa_run_batch")
        return [o[: out_bytes(b["n"], b["m"], self.flags(b.get("do_xor_key_rest", True),
                                                         b.get("do_compress", True)) & COMPRESS)]
                for o, b in zip(outs, blocks)]

    def batch_device(self, blocks, batch_flags=0, stream=None):
        """Device batch (pa_batch_device): blocks is a PaBlock array (or a list of PaBlock)."""
        if not isinstance(blocks, ctypes.Array):
            arr = (PaBlock * len(blocks))(*blocks)
        else:
            arr = blocks
        _check(self._L.pa_batch_device(self._h, arr, len(arr), batch_flags, stream), "pa_batch_device")

    def amplify_split(self, key, seed, n, m, sub_log2, do_xor_key_rest=True):
        """Sub-block (tiled, b = 2^sub_log2) evaluation of one block (pa_run_block_split)."""
        out = np.zeros(m // 8 + 8, dtype=np.uint8)
        k, kp = _u32(key)
        s, sp = _u32(seed)
        _check(self._L.pa_run_block_split(self._h, kp, sp, n, m, self.flags(do_xor_key_rest, True), sub_log2,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))), "pa_run_block_split")
        return out[: m // 8]

    def block_device_split(self, d_key, d_seed, n, m, flags, sub_log2, d_out, stream=None, d_margin=None):
        _check(self._L.pa_block_device_split(self._h, d_key, d_seed, n, m, flags, sub_log2, d_out, stream, d_margin),
               "pa_block_device_split")

    def host_empty(self, count, dtype=np.uint8):
        """A numpy array in pinned host memory (pa_host_alloc); freed when the array is collected.
        Blocks whose buffers are such arrays are copied by DMA without staging."""
        dtype = np.dtype(dtype)
        nbytes = max(1, int(count) * dtype.itemsize)
        p = ctypes.c_void_p()
        _check(self._L.pa_host_alloc(self._h, nbytes, ctypes.byref(p)), "pa_host_alloc")
        buf = (ctypes.c_uint8 * nbytes).from_address(p.value)
        arr = np.frombuffer(buf, dtype=dtype, count=int(count))
        L, h, addr = self._L, self._h.value, p.value
        self._pinned.append(weakref.finalize(buf, L.pa_host_free, h, addr))
        return arr

    def pinned_copy(self, a):
        out = self.host_empty(a.size, a.dtype)
        out[:] = a
        return out

    # ---- serving (continuous batching inside the engine)
    def serve_start(self, slots=3, max_batch=256, hbm_budget_bytes=16 << 30, batch_flags=0):
        cfg = PaServeConfig(slots, max_batch, hbm_budget_bytes, batch_flags, 0)
        _check(self._L.pa_serve_start(self._h, ctypes.byref(cfg)), "pa_serve_start")
        self._serve_keep = {}

    def serve_submit(self, key, n, m, out, seed=None, sid=0, do_xor_key_rest=True, do_compress=True):
        """Queues one block; `out` (uint8 array of at least out_bytes) receives amp_out.  The arrays
        are kept alive until serve_wait returns for the ticket."""
        b = PaBlock()
        k, _ = _u32(key)
        b.key = k.ctypes.data
        sd = None
        if seed is not None:
            sd, _ = _u32(seed)
            b.seed = sd.ctypes.data
        b.seed_id, b.n, b.m = sid, n, m
        b.flags = self.flags(do_xor_key_rest, do_compress)
        assert out.dtype == np.uint8 and out.flags.c_contiguous and out.size >= out_bytes(n, m, do_compress)
        b.out = out.ctypes.data
        t = ctypes.c_uint64()
        _check(self._L.pa_serve_submit(self._h, ctypes.byref(b), ctypes.byref(t)), "pa_serve_submit")
        self._serve_keep[t.value] = (k, sd, out)
        return t.value

    def serve_submit_many(self, reqs):
        """Queues several blocks under one lock.  reqs: (key, n, m, out, seed, sid) tuples, seed
        None for a resident seed sid; XOR and compression on.  Returns the tickets in order."""
        arr = (PaBlock * len(reqs))()
        keep = []
        for b, (key, n, m, out, seed, sid) in zip(arr, reqs):
            k, _ = _u32(key)
            b.key = k.ctypes.data
            sd = None
            if seed is not None:
                sd, _ = _u32(seed)
                b.seed = sd.ctypes.data
            b.seed_id, b.n, b.m, b.flags = sid, n, m, DEFAULT_FLAGS
            assert out.dtype == np.uint8 and out.flags.c_contiguous and out.size >= out_bytes(n, m, True)
            b.out = out.ctypes.data
            keep.append((k, sd, out))
        t = (ctypes.c_uint64 * len(reqs))()
        _check(self._L.pa_serve_submit_many(self._h, arr, len(reqs), t), "pa_serve_submit_many")
        for i, kp in enumerate(keep):
            self._serve_keep[t[i]] = kp
        return list(t)

    def serve_wait(self, ticket, timeout_ns=-1):
        """Returns the block's latency in nanoseconds (submit -> output written); raises on error."""
        lat = ctypes.c_uint64()
        st = self._L.pa_serve_wait(self._h, ticket, timeout_ns, ctypes.byref(lat))
        if st != -5:
            self._serve_keep.pop(ticket, None)
        _check(st, "pa_serve_wait")
        return lat.value

    def serve_stop(self):
        _check(self._L.pa_serve_stop(self._h), "pa_serve_stop")
        self._serve_keep = {}

    def selftest(self, name):
        """pa_selftest: (passed, report) for one of SELFTESTS (the reference's unitTest* modes)."""
        buf = ctypes.create_string_buffer(512)
        st = self._L.pa_selftest(self._h, name.encode(), buf, len(buf))
        if st not in (0, -8):
            _check(st, "pa_selftest")
        return st == 0, buf.value.decode()

    def stage_timing(self, enable=True):
        _check(self._L.pa_stage_timing(self._h, 1 if enable else 0), "pa_stage_timing")

    def stage_times(self):
        """{stage name: (total ms, launches)} since the last read (synchronises)."""
        cap = 32
        names = (ctypes.c_char_p * cap)()
        ms = (ctypes.c_double * cap)()
        cnt = (ctypes.c_uint32 * cap)()
        k = ctypes.c_uint32(cap)
        _check(self._L.pa_stage_times(self._h, names, ms, cnt, ctypes.byref(k)), "pa_stage_times")
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(k.value)}

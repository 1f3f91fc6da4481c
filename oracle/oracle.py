"""ctypes front of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg.  The product package never imports this module.
"""
import ctypes
import os
import pathlib
import subprocess

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"

DIRECT, NTT, FFT32, FFT64 = 0, 1, 2, 3
_lib = None


def build(force=False):
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not _LIB_PATH.exists():
        subprocess.run(["make", "-C", str(_HERE), "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
        u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
        L.pa_or_synth_bits.argtypes = [ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64, u32p]
        L.pa_or_key_split.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p]
        L.pa_or_bin_int2float.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, f32p]
        L.pa_or_bin_int2float.restype = ctypes.c_uint32
        L.pa_or_correction.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.pa_or_correction.restype = ctypes.c_float
        L.pa_or_element_wise_product.argtypes = [f32p, f32p, ctypes.c_uint32, ctypes.c_float]
        L.pa_or_to_binary_array.argtypes = [f32p, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                            ctypes.c_void_p, ctypes.c_int, u32p]
        L.pa_or_set_ntt_big.argtypes = [ctypes.c_int]
        L.pa_or_conv_bits.argtypes = [u32p, ctypes.c_uint32, u32p, ctypes.c_uint32, ctypes.c_uint32,
                                      u8p, ctypes.c_int]
        L.pa_or_conv_bits.restype = ctypes.c_int
        L.pa_or_amp_out.argtypes = [u32p, u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, u8p, ctypes.c_int]
        L.pa_or_amp_out.restype = ctypes.c_int64
        L.pa_or_fft32_seed_spectrum.argtypes = [u32p, ctypes.c_uint32, f32p, ctypes.POINTER(ctypes.c_uint32)]
        L.pa_or_fft32_seed_spectrum.restype = ctypes.c_int
        L.pa_or_amp_out_fft32_seeded.argtypes = [u32p, f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, u8p]
        L.pa_or_amp_out_fft32_seeded.restype = ctypes.c_int64
        L.pa_or_fft_values.argtypes = [u32p, ctypes.c_uint32, u32p, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_int, f64p]
        L.pa_or_fft_values.restype = ctypes.c_int
        L.pa_or_conv_counts.argtypes = [u32p, ctypes.c_uint32, u32p, ctypes.c_uint32, ctypes.c_uint32,
                                        u32p]
        L.pa_or_conv_counts.restype = ctypes.c_int
        L.pa_or_set_threads.argtypes = [ctypes.c_int]
        L.pa_or_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def set_threads(n):
    lib().pa_or_set_threads(int(n))


def get_threads():
    return int(lib().pa_or_get_threads())


def synth_bits(seed, nbits, p=0.5):
    out = np.zeros((nbits + 31) // 32, dtype=np.uint32)
    lib().pa_or_synth_bits(seed, p, nbits, out)
    return out


def key_split(key, n, m):
    ks = np.zeros(n // 32, dtype=np.uint32)
    kr = np.zeros(max(m // 32, 1), dtype=np.uint32)
    lib().pa_or_key_split(np.ascontiguousarray(key, dtype=np.uint32), n, m, ks, kr)
    return ks, kr[: m // 32]


def bin_int2float(bits, nbits, reduction_exp=11):
    out = np.zeros(nbits, dtype=np.float32)
    ones = lib().pa_or_bin_int2float(np.ascontiguousarray(bits, dtype=np.uint32), nbits,
                                     reduction_exp, out)
    return out, int(ones)


def correction(wk, ws, n):
    return float(lib().pa_or_correction(wk, ws, n))


def element_wise_product(a, b, r):
    a = np.ascontiguousarray(a, dtype=np.float32).copy()
    lib().pa_or_element_wise_product(a, np.ascontiguousarray(b, dtype=np.float32), a.size // 2, r)
    return a


def to_binary_array(v, nbits, norm, corr, key_rest=None, reverse_endian=True):
    v = np.ascontiguousarray(v, dtype=np.float32)
    out = np.zeros((nbits + 31) // 32, dtype=np.uint32)
    kr = None
    if key_rest is not None:
        kr = np.ascontiguousarray(key_rest, dtype=np.uint32)
    lib().pa_or_to_binary_array(v, nbits, norm, corr,
                                kr.ctypes.data if kr is not None else None,
                                1 if reverse_endian else 0, out)
    return out


def conv_bits(key_start, kbits, seed, n, nout, method=DIRECT):
    out = np.zeros((nout + 7) // 8, dtype=np.uint8)
    rc = lib().pa_or_conv_bits(np.ascontiguousarray(key_start, dtype=np.uint32), kbits,
                               np.ascontiguousarray(seed, dtype=np.uint32), n, nout, out, method)
    if rc != 0:
        raise ValueError(f"pa_or_conv_bits failed (rc={rc})")
    return out


def conv_counts(key_start, kbits, seed, n, nout):
    out = np.zeros(nout, dtype=np.uint32)
    rc = lib().pa_or_conv_counts(np.ascontiguousarray(key_start, dtype=np.uint32), kbits,
                                 np.ascontiguousarray(seed, dtype=np.uint32), n, nout, out)
    if rc != 0:
        raise ValueError("pa_or_conv_counts failed")
    return out


def fft_values(key_start, kbits, seed, n, nout, use_double=False):
    out = np.zeros(nout, dtype=np.float64)
    rc = lib().pa_or_fft_values(np.ascontiguousarray(key_start, dtype=np.uint32), kbits,
                                np.ascontiguousarray(seed, dtype=np.uint32), n, nout,
                                1 if use_double else 0, out)
    if rc != 0:
        raise ValueError("pa_or_fft_values failed")
    return out


def amp_out(key, seed, n, m, do_xor=True, do_compress=True, reverse_endian=True, method=NTT):
    """One block with the reference's wire semantics; returns the output bytes."""
    out_bits = m if do_compress else n // 2
    nbytes = ((out_bits + 31) // 32) * 4
    out = np.zeros(nbytes, dtype=np.uint8)
    rc = lib().pa_or_amp_out(np.ascontiguousarray(key, dtype=np.uint32),
                             np.ascontiguousarray(seed, dtype=np.uint32), n, m, int(do_xor),
                             int(do_compress), int(reverse_endian), out, method)
    if rc < 0:
        raise ValueError("pa_or_amp_out: bad arguments")
    return out[: (out_bits + 7) // 8] if reverse_endian else out[: ((out_bits + 31) // 32) * 4]


def fft32_seed_spectrum(seed, n):
    """The static-seed FP32 restatement's per-seed half: (S, ws) for amp_out_fft32_seeded."""
    S = np.zeros(n + 2, dtype=np.float32)
    ws = ctypes.c_uint32()
    if lib().pa_or_fft32_seed_spectrum(np.ascontiguousarray(seed, dtype=np.uint32), n, S, ctypes.byref(ws)) != 0:
        raise MemoryError("pa_or_fft32_seed_spectrum")
    return S, ws.value


def amp_out_fft32_seeded(key, spectrum, n, m, do_xor=True, do_compress=True, reverse_endian=True):
    """One block of the FP32 restatement with a precomputed seed spectrum (static seed)."""
    S, ws = spectrum
    out_bits = m if do_compress else n // 2
    out = np.zeros(((out_bits + 31) // 32) * 4, dtype=np.uint8)
    rc = lib().pa_or_amp_out_fft32_seeded(np.ascontiguousarray(key, dtype=np.uint32), S, ws, n, m, int(do_xor),
                                          int(do_compress), int(reverse_endian), out)
    if rc < 0:
        raise ValueError("pa_or_amp_out_fft32_seeded: bad arguments")
    return out[: (out_bits + 7) // 8] if reverse_endian else out


def set_ntt_big(on):
    """Test hook: route every NTT through the long-length prime 3*2^30+1."""
    lib().pa_or_set_ntt_big(1 if on else 0)

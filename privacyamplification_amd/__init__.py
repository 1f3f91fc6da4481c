"""MI355X (gfx950) Toeplitz-hashing engine for privacy amplification.

A drop-in for the Toeplitz-hashing stage of nicoboss/PrivacyAmplification: the same block contract
(key / seed / amp_out bit layouts, vertical_len, do_xor_key_rest, do_compress) computed by
hand-written HIP kernels behind the C ABI declared in include/pa_engine.h.
"""
from .engine import (  # noqa: F401
    BATCH_GRAPH,
    BUFFERS,
    BATCH_TUNE,
    COMPRESS,
    DEFAULT_FLAGS,
    XOR_KEY_REST,
    Engine,
    EngineError,
    PaBlock,
    PaServeConfig,
    SELFTESTS,
    block_extents,
    build,
    library_path,
    load_library,
    out_bytes,
    synth_bits,
    workspace_bytes,
)

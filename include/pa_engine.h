/*
 * pa_engine.h -- C ABI of the MI355X Toeplitz-hashing engine (privacy amplification).
 *
 * Drop-in for the Toeplitz-hashing stage of nicoboss/PrivacyAmplification.  One call computes,
 * for a key block of n+1 bits and an n-bit Toeplitz seed, the m-bit
 *
 *     amp_out[p] = XOR_{j=0..h} key[j] * seed[(p - j) mod n]   XOR   key[h + 1 + p],   h = n - m,
 *
 * bit-exactly as the reference's mainloop does (PrivacyAmplification.cu:1819-2042), on gfx950
 * through hand-written HIP kernels.  All pointers are plain host or device pointers; no C++ or
 * torch types cross this boundary.  Every entry point returns PA_OK (0) or a negative pa_status.
 *
 * Bit layouts are the reference's wire formats:
 *   key / seed (input) : RefBits -- little-endian uint32 words, bit p of the stream is
 *                        (w[p>>5] >> (31-(p&31))) & 1   (PrivacyAmplification.cu:584-593).
 *   amp_out (output)   : OutBytes -- byte p>>3, bit 7-(p&7), the AMPOUT_REVERSE_ENDIAN=TRUE layout
 *                        (PrivacyAmplification.h:38, .cu:260-267); with out_reverse_endian=0 the
 *                        output is RefBits words instead.
 * Key buffer length: n/32+1 words with PA_XOR_KEY_REST (the reference's key_blocks,
 * PrivacyAmplification.cu:1428), h/32+1 words without it (reciveDataKey's NoXOR frame,
 * PrivacyAmplification.cu:1068-1072).  Seed: n/32 words (desired_block).
 * Output length: m/8 bytes with PA_COMPRESS, n/16 bytes without (sendData,
 * PrivacyAmplification.cu:1186).
 */
#ifndef PA_ENGINE_H
#define PA_ENGINE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct pa_engine pa_engine;
typedef uint64_t pa_seed_id;
typedef uint64_t pa_ticket;

/* Block flags: the two 1-byte frames the key server sends ahead of each key
 * (do_xor_key_rest, do_compress; PrivacyAmplification.cu:1055-1056). */
enum pa_flags {
    PA_XOR_KEY_REST = 1u,
    PA_COMPRESS = 2u,
    PA_DEFAULT_FLAGS = 3u
};

enum pa_status {
    PA_OK = 0,
    PA_ERR_ARG = -1,         /* null pointer, bad flag, unknown handle */
    PA_ERR_SIZE = -2,        /* n / m outside the supported family (the reference exits 400) */
    PA_ERR_DEVICE = -3,      /* HIP runtime error */
    PA_ERR_NOMEM = -4,       /* device allocation failed or HBM budget exceeded */
    PA_ERR_TIMEOUT = -5,     /* pa_wait timed out */
    PA_ERR_NOT_FOUND = -6,   /* seed id / ticket unknown */
    PA_ERR_UNSUPPORTED = -7, /* valid request this build does not serve */
    PA_ERR_CHECK = -8        /* a self-test (pa_selftest) found a wrong result */
};

/* Engine configuration.  The fields mirror config.yaml (PrivacyAmplification/config.yaml:6-124,
 * defaults PrivacyAmplification.cu:1244-1271) where a field exists there. */
typedef struct {
    int32_t device;                 /* gpu_device_id_to_use */
    uint32_t max_n_log2;            /* largest factor_exp accepted (the reference's maximum is 27) */
    uint32_t reduction_exp;         /* reduction_exp (config.yaml:42); scaling only, parity-neutral */
    uint32_t pre_mul_reduction_exp; /* pre_mul_reduction_exp (config.yaml:47); idem */
    uint64_t hbm_budget_bytes;      /* admission budget for resident seeds + in-flight blocks */
    uint32_t out_reverse_endian;    /* AMPOUT_REVERSE_ENDIAN (PrivacyAmplification.h:38), default 1 */
    uint32_t reserved;
} pa_config;

/* Fills the defaults: device 0, max_n_log2 27, reduction 11 / 5, budget 64 GiB, reverse endian. */
int pa_config_default(pa_config* cfg);

/* Creates an engine on cfg->device (main's device set-up and buffer allocation,
 * PrivacyAmplification.cu:1440-1620).  Device buffers are allocated lazily per block size. */
int pa_create(const pa_config* cfg, pa_engine** out);
int pa_destroy(pa_engine* eng);

/* Registers a resident Toeplitz seed: its spectrum is computed once and kept in HBM, the
 * MI355X form of a reused seed (reuse_seed_amount != 0, PrivacyAmplification.cu:1781-1791,
 * 1898-1910).  seed: n/32 words of host memory.  reuse is kept as metadata (-1 forever, 0 one
 * block, k>0 k blocks); the caller drops the seed with pa_seed_drop.  m fixes the block shape
 * the spectrum serves (the key split depends on it only for non-power-of-two n). */
int pa_seed_put(pa_engine* eng, const uint32_t* seed, uint32_t n, uint32_t m, int32_t reuse,
                pa_seed_id* id);
int pa_seed_drop(pa_engine* eng, pa_seed_id id);

/* One block, synchronous, host buffers (the body of one mainloop iteration together with the
 * receive-side key split and the send-side output, PrivacyAmplification.cu:861-892, 1819-2042,
 * 1143-1225).  seed: n/32 words (dynamic seed). */
int pa_run_block(pa_engine* eng, const uint32_t* key, const uint32_t* seed, uint32_t n,
                 uint32_t m, uint32_t flags, uint8_t* out);

/* Same with a resident seed (static-seed mode). */
int pa_run_block_seeded(pa_engine* eng, const uint32_t* key, pa_seed_id seed, uint32_t n,
                        uint32_t m, uint32_t flags, uint8_t* out);

/* Asynchronous form: enqueues one block on the engine's stream and returns a ticket; the host
 * buffers must stay valid until pa_wait returns for that ticket.  seed_id == 0 with seed != NULL
 * runs a dynamic seed; seed == NULL uses the resident seed seed_id. */
int pa_submit(pa_engine* eng, const uint32_t* key, const uint32_t* seed, pa_seed_id seed_id,
              uint32_t n, uint32_t m, uint32_t flags, uint8_t* out, pa_ticket* ticket);
/* timeout_ns < 0 waits forever.  One waiting thread per ticket: a second pa_wait on a ticket
 * that another thread is waiting on returns PA_ERR_ARG; after success the ticket is gone
 * (PA_ERR_NOT_FOUND).  A timed-out ticket may be waited on again. */
int pa_wait(pa_engine* eng, pa_ticket ticket, int64_t timeout_ns);

/* Device-resident form: key, seed and out are device pointers, the work is stream-ordered on
 * `stream` (a hipStream_t, NULL = the engine's stream) and nothing is synchronised.  The engine's
 * per-size workspace is shared: a call on another stream than the previous call of the same
 * transform length first waits (on the device) for that call's kernels, so calls on several
 * streams serialise per size.  Inside a graph capture on `stream` no such wait is recorded.
 * margin_out (device float, may be NULL) receives max |v - rint(v)| over the block's pre-rounding
 * values (atomic max; the caller zeroes it). */
int pa_block_device(pa_engine* eng, const uint32_t* d_key, const uint32_t* d_seed, uint32_t n,
                    uint32_t m, uint32_t flags, uint8_t* d_out, void* stream, float* d_margin);
int pa_block_device_seeded(pa_engine* eng, const uint32_t* d_key, pa_seed_id seed, uint32_t n,
                           uint32_t m, uint32_t flags, uint8_t* d_out, void* stream,
                           float* d_margin);

/* pa_block_device that allocates nothing.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  The block's scratch (spectra, bit masks, nibbles, popcount partials)
 * lives in one file-scope device array of 64 MiB per GPU, held until the process ends; the twiddle
 * tables it needs are made by pa_create.  Serves transform lengths up to 2^22 bits (n <= 2^22);
 * larger blocks return PA_ERR_SIZE (pa_strerror names the limit).  Same arguments and output as
 * pa_block_device.  Every call on a GPU shares that GPU's array and leaves its contents behind:
 * calls that overlap on one GPU (other streams, other engines, other threads) race. */
int pa_block_device_noalloc(pa_engine* eng, const uint32_t* d_key, const uint32_t* d_seed, uint32_t n,
                            uint32_t m, uint32_t flags, uint8_t* d_out, void* stream, float* d_margin);

/* One block of a batch.  For pa_batch_device every pointer is a device pointer; for pa_run_batch
 * key, seed and out are host pointers and margin must be NULL. */
typedef struct {
    const uint32_t* key;   /* RefBits key: n/32+1 words (PA_XOR_KEY_REST) or h/32+1 words */
    const uint32_t* seed;  /* n/32 seed words, or NULL to use the resident seed seed_id */
    pa_seed_id seed_id;
    uint32_t n, m, flags;
    uint32_t reserved;     /* 0 */
    uint8_t* out;          /* m/8 bytes (PA_COMPRESS) or n/16 bytes */
    float* margin;         /* device float receiving max |v - rint(v)| (atomic max), or NULL */
} pa_block;

/* Batch options: PA_BATCH_GRAPH replays a captured hipGraph for a batch of the same shape
 * (lengths classes, seed modes and counts); PA_BATCH_TUNE times the launch shapes of a shape not
 * seen before (pass A walk length, then row-pass grid) and keeps the fastest. */
enum pa_batch_flags {
    PA_BATCH_GRAPH = 1u,
    PA_BATCH_TUNE = 2u
};

/* Many blocks of uneven length in few launches: blocks are grouped by transform length and seed
 * mode; each group runs as one launch per pass (grid z = block), and groups run concurrently on
 * the engine's fork streams.  The batched form of the reference's bank of 2^27/n equal blocks per
 * batched cuFFT plan (PrivacyAmplification_subblocks.cu:1058-1081, 1254-1403), here with any mix
 * of n, m, flags and static/dynamic seeds.  Stream-ordered on `stream`, nothing synchronised.
 * At most 65535 blocks per call. */
int pa_batch_device(pa_engine* eng, const pa_block* blocks, uint32_t count, uint32_t batch_flags,
                    void* stream);
/* Same with host buffers; returns when every output has been written. */
int pa_run_batch(pa_engine* eng, const pa_block* blocks, uint32_t count, uint32_t batch_flags);

/* Sub-block (tiled) evaluation of one block, the device form of LargeBlocksizeExample's tiling
 * (LBS/main.cpp:104-124, 313-523) used as the precision fallback: the m x (h+1) Toeplitz product is
 * cut into b x b tiles, b = 2^sub_log2; each tile is a length-2b circular product through the
 * batched engine (the LBS tile contract: NoXOR, no compression), tiles XOR into their output band,
 * then the key rest is XORed.  Every convolution value stays below b, so FP32 keeps a wide rounding
 * margin at any n; the cost grows as n^2 / b.  Output identical to pa_block_device.  Requires
 * PA_COMPRESS and 2^6 <= b < n, 2b <= 2^max_n_log2; n, m multiples of 32 (n < 2^31).
 * d_margin as in pa_block_device (max over all tiles). */
int pa_block_device_split(pa_engine* eng, const uint32_t* d_key, const uint32_t* d_seed, uint32_t n,
                          uint32_t m, uint32_t flags, uint32_t sub_log2, uint8_t* d_out, void* stream,
                          float* d_margin);
/* Same with host buffers (synchronous). */
int pa_run_block_split(pa_engine* eng, const uint32_t* key, const uint32_t* seed, uint32_t n,
                       uint32_t m, uint32_t flags, uint32_t sub_log2, uint8_t* out);

/* Pinned host memory for key, seed and output buffers: the counterpart of the reference's
 * cudaMallocHost'd input and output block caches (PrivacyAmplification.cu:1535-1551).  The memory
 * is mapped into the device: blocks whose buffers lie in it are read and written there directly by
 * the engine's copy kernels; other host memory is staged through the engine's own pinned buffers. */
int pa_host_alloc(pa_engine* eng, uint64_t bytes, void** ptr);
int pa_host_free(pa_engine* eng, void* ptr);

/* Serving: a scheduler thread inside the engine admits submitted blocks into batches while
 * earlier batches are still on the GPU (continuous batching), within an HBM budget for in-flight
 * blocks.  The native form of the reference's receive -> mainloop -> send pipeline with its input
 * and output block caches (input_blocks_to_cache / output_blocks_to_cache, spin-wait rings,
 * PrivacyAmplification.cu:202-207, 1686-1717, 2098-2103). */
typedef struct {
    uint32_t slots;            /* batches in flight at once (default 3) */
    uint32_t max_batch;        /* blocks per batch (default 256) */
    uint64_t hbm_budget_bytes; /* device memory for in-flight blocks, split evenly over the slots
                                  (default 16 GiB); a block larger than one slot's share fails
                                  with PA_ERR_NOMEM.  pa_serve_start reserves a quarter of each
                                  share (at most 4 GiB, plus half again) up front */
    uint32_t batch_flags;      /* PA_BATCH_* for every batch (default 0) */
    uint32_t reserved;
} pa_serve_config;

/* Starts the scheduler (cfg NULL: defaults).  At most one server per engine. */
int pa_serve_start(pa_engine* eng, const pa_serve_config* cfg);
/* Queues one block (host pointers, which must stay valid until pa_serve_wait returns for it);
 * returns at once with a ticket.  Resident seeds must not be dropped while blocks use them. */
int pa_serve_submit(pa_engine* eng, const pa_block* block, pa_ticket* ticket);
/* Queues count blocks under one lock, tickets[i] for blocks[i]: for a caller that holds several
 * blocks at once (a filled input cache).  Every block is checked first; on an error nothing is
 * queued. */
int pa_serve_submit_many(pa_engine* eng, const pa_block* blocks, uint32_t count, pa_ticket* tickets);
/* Waits for a block (timeout_ns < 0: forever).  On success *latency_ns (may be NULL) receives the
 * time from pa_serve_submit to the output being written; returns the block's own status.  One
 * waiting thread per ticket (a second concurrent wait returns PA_ERR_ARG). */
int pa_serve_wait(pa_engine* eng, pa_ticket ticket, int64_t timeout_ns, uint64_t* latency_ns);
/* Drains every queued and in-flight block, lets the threads already inside pa_serve_wait return,
 * then stops the scheduler.  No pa_serve_* call may start once pa_serve_stop has begun.
 * pa_destroy calls it first when a server is running. */
int pa_serve_stop(pa_engine* eng);

/* Per-stage device timing, the counterpart of the reference's STOPWATCH stage timers
 * (PrivacyAmplification.cu:146-166, 2053-2074).  When enabled, the engine records HIP events
 * around every kernel it launches (on the launch stream).  pa_stage_times synchronises, then
 * returns per stage: its name (static string), the summed milliseconds and the launch count
 * since the last read, and resets the sums.  *n_stages is in: capacity, out: stages written. */
int pa_stage_timing(pa_engine* eng, int enable);
int pa_stage_times(pa_engine* eng, const char** names, double* ms, uint32_t* counts,
                   uint32_t* n_stages);

/* Bytes of device workspace a block of this shape uses (0 for an unsupported shape). */
uint64_t pa_workspace_bytes(uint32_t n, uint32_t m, uint32_t flags);

/* The device buffers of one block. */
enum pa_buffer {
    PA_BUF_KEY = 0,        /* key words (caller's, or the engine's staging of them) */
    PA_BUF_SEED,           /* seed words */
    PA_BUF_OUT,            /* output bytes */
    PA_BUF_SPEC_KEY,       /* key spectrum, then the product (four-step path) */
    PA_BUF_SPEC_SEED,      /* seed spectrum */
    PA_BUF_BITS_KEY,       /* key bit masks of pass A */
    PA_BUF_BITS_SEED,      /* seed bit masks */
    PA_BUF_NIB,            /* rounded output bits before packing */
    PA_BUF_COUNTS,         /* popcount partials */
    PA_BUF_TILES,          /* pass B's output in tiles for pass A' (transform lengths 2^24..2^27) */
    PA_BUF_COUNT
};
/* Buffer extents of a block shape, a diagnostic with no counterpart in the reference: touched[b]
 * is the number of bytes of buffer b that the block's kernels index at most, alloc[b] the bytes
 * every allocation path gives it (the smallest over the per-size workspace, pa_submit, batch and
 * serving staging; for key, seed and out also the caller contract of the device entry points:
 * key_words * 4, n/8 and m/8 or n/16 bytes).  The engine refuses a shape with touched > alloc
 * (PA_ERR_UNSUPPORTED).  *transform_len receives L.  Runs on the host only. */
int pa_block_extents(uint32_t n, uint32_t m, uint32_t flags, uint32_t* transform_len,
                     uint64_t touched[PA_BUF_COUNT], uint64_t alloc[PA_BUF_COUNT]);

/* Synthetic RefBits: xoshiro256** seeded through splitmix64, Bernoulli(p) bits; p == 0.5 takes
 * two words per 64-bit draw (high half first).  Host memory, nwords = ceil(nbits/32). */
int pa_synth_bits(uint64_t seed, double p, uint64_t nbits, uint32_t* out);

/* Device self-tests, the counterpart of the reference's unitTest* argv modes
 * (PrivacyAmplification.cu:299-737, dispatched at :1645-1652).  name is the reference's test name
 * without "unitTest": "CalculateCorrectionFloat", "SetFirstElementToZero", "ElementWiseProduct",
 * "BinInt2float" or "ToBinaryArray".  Each runs this engine's device code for that step (its device
 * functions or its production kernels) on the reference test's closed-form inputs and checks the
 * result on the host.  report (may be NULL) receives a one-line summary or the first mismatch.
 * Returns PA_OK when the test passed, PA_ERR_CHECK when it found a wrong value, PA_ERR_ARG for an
 * unknown name. */
int pa_selftest(pa_engine* eng, const char* name, char* report, uint32_t report_bytes);

const char* pa_strerror(int status);
const char* pa_version(void);

#ifdef __cplusplus
}
#endif
#endif

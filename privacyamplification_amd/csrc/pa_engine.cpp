// pa_engine.cpp -- host side of the MI355X Toeplitz-hashing engine: the C ABI of
// include/pa_engine.h, block planning, device workspaces, twiddle tables and resident seeds.
//
// The reference runs one block per mainloop iteration with a synchronise after every step
// (PrivacyAmplification.cu:1686-2118) and re-plans its FFTs on every mainloop entry
// (:1712-1756).  Here a block is a short stream-ordered kernel sequence on the engine's stream;
// workspaces and twiddle tables are created once per transform length and kept.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pa_engine.h"
#include "pa_geom.hpp"
#include "pa_kernels.hpp"

namespace {

using pa::Geom;

#define PA_HIP(expr)                                              \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) {                                   \
            set_error(e_, #expr, __FILE__, __LINE__);             \
            return PA_ERR_DEVICE;                                 \
        }                                                         \
    } while (0)

thread_local char g_last_error[512] = "";
thread_local char g_size_error[256] = "";  // detail of the last PA_ERR_SIZE that has one
void set_error(hipError_t e, const char* what, const char* file, int line) {
    snprintf(g_last_error, sizeof(g_last_error), "%s: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

bool is_pow2(uint32_t x) { return x && !(x & (x - 1)); }
uint32_t log2u(uint64_t x) {
    uint32_t r = 0;
    while ((1ull << r) < x) ++r;
    return r;
}

// Plans the transform for a block: see pa_geom.hpp.  Returns PA_OK or an error status.
int make_geom(uint32_t n, uint32_t m, uint32_t flags, uint32_t reverse_endian, uint32_t max_n_log2,
              Geom* g) {
    g_size_error[0] = '\0';
    if (flags & ~3u) return PA_ERR_ARG;
    if (n < 64 || (n % 32) || m == 0 || (m % 32) || m >= n) return PA_ERR_SIZE;
    const bool pow2 = is_pow2(n);
    const bool compress = (flags & PA_COMPRESS) != 0;
    if (!pow2 && !compress) return PA_ERR_UNSUPPORTED;  // n/2-bit output is defined for 2^k only
    const uint32_t L = pow2 ? n : (1u << log2u(n));
    if (log2u(L) > pa::MAX_L_LOG || log2u(L) > max_n_log2) return PA_ERR_SIZE;
    Geom x{};
    x.n = n;
    x.m = m;
    x.h = n - m;
    x.kbits = x.h + 1;
    x.L = L;
    x.N = L / 2;
    x.lgN = log2u(x.N);
    if (x.lgN <= (uint32_t)pa::SMALL_MAX_LOG_N) {
        x.N1 = 1;
        x.N2 = x.N;
        x.lg1 = 0;
        x.lg2 = x.lgN;
        x.M = 1;
        x.D = 1;
        x.W = 1;
    } else {
        x.lg2 = pa::ROW_LOG;
        x.N2 = 1u << x.lg2;
        x.lg1 = x.lgN - x.lg2;
        x.N1 = 1u << x.lg1;
        x.M = x.N1 <= 512 ? x.N1 : (x.N1 / 16 > 512 ? x.N1 / 16 : 512);  // D <= 16 bit classes
        x.D = x.N1 / x.M;
        x.W = (uint32_t)pa::inv_cols_width((int)x.lg1);
    }
    x.out_bits = compress ? m : n / 2;
    x.out_off = pow2 ? 0 : x.h;
    x.seed_words = n / 32;
    x.seed_rot = pow2 ? 0 : m / 32;
    x.key_hw = x.h / 32;
    x.do_xor = (flags & PA_XOR_KEY_REST) ? 1u : 0u;
    x.reverse_endian = reverse_endian ? 1u : 0u;
    x.inv_N = 1.0f / (8.0f * (float)x.N);  // exact: includes the factor 8 the real splits leave
    *g = x;
    return PA_OK;
}

uint32_t key_words_for(const Geom& g) { return g.do_xor ? g.n / 32 + 1 : g.key_hw + 1; }

// ------------------------------------------------------------------ buffer extents ----
// Every device buffer of one block: the bytes its kernels index at most (`touched`, from the
// kernels' own index arithmetic) and the bytes every allocation path gives it (`alloc`, the
// smallest over the workspace, pa_submit, batch, serving and split paths and the caller contract
// of the device entry points).  All allocation sites take their sizes from the stage_* functions
// below, and prepare() refuses a shape whose touched bytes exceed its allocation.
size_t out_bytes_of(const Geom& g) { return g.out_bits / 8; }
// engine staging of one key: n/32 + 2 words (the kernels read n/32 + 1 at most; one zero word
// more keeps a NoXOR key's staging defined past its h/32 + 1 words)
uint64_t stage_key_bytes(const Geom& g) { return (uint64_t)(g.n / 32 + 2) * 4; }
uint64_t stage_seed_bytes(const Geom& g) { return (uint64_t)(g.n / 32) * 4; }
// engine staging of one output: the output plus 16 bytes
uint64_t stage_out_bytes(const Geom& g) { return out_bytes_of(g) + 16; }
// a length-L workspace's staging, shared by every block of that length (n <= L; the output, m/8
// bytes with compression and n/16 without, m < n, plus 16 fits in L/8 + 16)
uint64_t ws_key_bytes(uint32_t L) { return (uint64_t)(L / 32 + 2) * 4; }
uint64_t ws_out_bytes(uint32_t L) { return (uint64_t)L / 8 + 16; }

struct Extents {
    uint64_t touched[PA_BUF_COUNT] = {};
    uint64_t alloc[PA_BUF_COUNT] = {};
    bool ok() const {
        for (int i = 0; i < PA_BUF_COUNT; ++i)
            if (touched[i] > alloc[i]) return false;
        return true;
    }
};

Extents block_extents(const Geom& g) {
    Extents x;
    const uint64_t N = g.N;
    // key: k_transpose_bits / k_small read words [0, h/32] through key_word(); with the key-rest
    // XOR, k_pack / k_small read key[h/32 + w + 1] for 32 w < m, i.e. up to word n/32
    x.touched[PA_BUF_KEY] = (uint64_t)(g.do_xor ? g.n / 32 + 1 : g.key_hw + 1) * 4;
    x.touched[PA_BUF_SEED] = (uint64_t)g.seed_words * 4;  // seed_word(): words [0, n/32)
    // k_pack (four-step) and k_small write output words [0, out_bits / 32)
    x.touched[PA_BUF_OUT] = (uint64_t)g.out_bits / 8;
    x.alloc[PA_BUF_KEY] = std::min(stage_key_bytes(g), ws_key_bytes(g.L));
    x.alloc[PA_BUF_SEED] = std::min(stage_seed_bytes(g), ws_key_bytes(g.L));
    // the device entry points' caller contract: m/8 bytes (compress) or n/16; staging is larger
    x.alloc[PA_BUF_OUT] = std::min({(uint64_t)out_bytes_of(g), stage_out_bytes(g), ws_out_bytes(g.L)});
    x.alloc[PA_BUF_COUNTS] = 2 * pa::POP_SLOTS * 4;
    if (g.N1 == 1) return x;  // the one-workgroup path keeps everything else on chip
    x.touched[PA_BUF_COUNTS] = 2 * pa::POP_SLOTS * 4;
    // k_transpose_bits writes tb[u N2 + t2] for u < M, t2 < N2
    x.touched[PA_BUF_BITS_KEY] = x.touched[PA_BUF_BITS_SEED] = (uint64_t)g.M * g.N2 * 4;
    // pass A writes Y[k1 N2 + t2], k1 < N1; the seed-spectrum pass adds the Nyquist value S[N]
    x.touched[PA_BUF_SPEC_KEY] = N * 8;
    x.touched[PA_BUF_SPEC_SEED] = (N + 1) * 8;
    // pass A' writes crumbs [0, N) (one column per workgroup) or nibbles [0, N/2); k_pack reads
    // 4 bytes at (column or pair) * N1 + row for rows row < N1, row a multiple of 4
    {
        const uint64_t cols = g.W == 1 ? g.N2 : g.N2 / 2;
        const uint64_t last_row = g.N1 >= 4 ? g.N1 - 4 : 0;
        const uint64_t wr = cols * g.N1, rd = (cols - 1) * g.N1 + last_row + 4;
        x.touched[PA_BUF_NIB] = std::max(wr, rd);
    }
    x.alloc[PA_BUF_SPEC_KEY] = N * 8;
    x.alloc[PA_BUF_SPEC_SEED] = (N + 1) * 8;
    x.alloc[PA_BUF_BITS_KEY] = x.alloc[PA_BUF_BITS_SEED] = (uint64_t)(N / g.D) * 4;
    x.alloc[PA_BUF_NIB] = pa::nib_bytes(g.N, g.W);
    if (pa::rt_tiled(g.lg1)) {
        // pass B writes tile ((t2 / RT_COLS) N1/2 + p) 2 RT_COLS + s RT_COLS + t2 mod RT_COLS for
        // t2 < N2, p < N1/2, s < 2 (the last: N - 1); pass A' reads the same range
        const uint64_t last = ((uint64_t)(g.N2 - 1) / pa::RT_COLS * (g.N1 / 2) + (g.N1 / 2 - 1)) * (2 * pa::RT_COLS) +
                              pa::RT_COLS + (g.N2 - 1) % pa::RT_COLS;
        x.touched[PA_BUF_TILES] = (last + 1) * 8;
        x.alloc[PA_BUF_TILES] = N * 8;
    }
    return x;
}

// Per transform length: w_L^x = lo[x mod 2^14] * hi[x >> 14].
struct Twiddles {
    float2* lo = nullptr;      // 2^14 entries
    float2* hi = nullptr;      // max(1, L / 2^14) entries
};

// Per transform length: the device scratch of one (unbatched) block.
struct Workspace {
    uint32_t L = 0;
    float2* Yk = nullptr;      // N complex (key rows / R')
    float2* Ys = nullptr;      // N+1 complex (seed rows / split seed spectrum)
    float2* Rt = nullptr;      // N complex: pass B's output tiles (rt_tiled(lg1) only)
    uint32_t* tbk = nullptr;   // N/D words
    uint32_t* tbs = nullptr;
    uint8_t* nib = nullptr;    // L/4 bytes
    uint32_t* counts = nullptr;  // POP_SLOTS key popcount partials, then POP_SLOTS seed partials
    bool counts_zero = false;    // counts are known to be zero (the last block's pack cleared them)
    uint32_t* d_key = nullptr;   // staging for the host-buffer API
    uint32_t* d_seed = nullptr;
    uint32_t* d_out = nullptr;
    uint64_t bytes = 0;
    // The last stream that used this workspace and an event recorded after its last kernel: a
    // block on another stream waits for it, so calls on two streams serialise instead of
    // overwriting each other's spectra.
    hipStream_t last_stream = nullptr;
    hipEvent_t last_done = nullptr;
};

struct Seed {
    Geom g{};
    float2* spec = nullptr;    // N+1 complex split spectrum
    uint32_t* ws = nullptr;    // device popcount (POP_SLOTS partials)
    int32_t reuse = 0;
    uint64_t bytes = 0;
};

struct Ticket {
    hipEvent_t done = nullptr;
    uint8_t* host_out = nullptr;
    uint32_t* dev_out = nullptr;  // owned output staging
    uint32_t* dev_key = nullptr;
    uint32_t* dev_seed = nullptr;
    size_t out_bytes = 0;
    bool waiting = false;  // a pa_wait call is polling this ticket's event
};

// A batch in flight: its device arena (block table, scratch, staging), the pinned host copy of its
// block table, fork streams for concurrent groups, and the graphs captured for its shapes.
struct BatchSlot {
    static constexpr int FORKS = 4;
    hipStream_t stream = nullptr;              // the slot's own launch stream (serving)
    hipStream_t fork[FORKS] = {};
    hipEvent_t fork_ev = nullptr, join_ev[FORKS] = {};
    uint8_t* arena = nullptr;                  // grow-only device arena
    uint64_t arena_cap = 0;
    pa::BlockArgs* htab = nullptr;             // pinned, mapped host copy of the block table
    uint8_t* htab_dev = nullptr;               // its device-mapped address
    uint32_t htab_cap = 0;
    pa::CopyDesc* hdesc = nullptr;             // staging copy descriptors (pinned, mapped)
    const pa::CopyDesc* hdesc_dev = nullptr;
    uint32_t hdesc_cap = 0;
    hipEvent_t htab_free = nullptr;            // recorded after the last read of htab / hdesc
    bool htab_pending = false;
    std::map<std::string, hipGraphExec_t> graphs;
};

// ------------------------------------------------------------------ allocation-free scratch ----
// pa_block_device_noalloc's device scratch: one array per GPU, held until the process ends.  In
// host code `scratch` is only a stand-in for the symbol; its device address comes from
// hipGetSymbolAddress, made once per device at pa_create with that device current, and only that
// address is ever handed to kernels.
constexpr uint32_t NOALLOC_MAX_L_LOG = 22;
constexpr uint64_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device `dev` (looked up on first use).
hipError_t scratch_address(int dev, void** out) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t r = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (r != hipSuccess) return r;
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

struct pa_engine {
    pa_config cfg{};
    hipStream_t stream = nullptr;
    hipStream_t upload = nullptr;        // pa_run_block: a fresh seed's host-to-device copy
    hipEvent_t upload_done = nullptr, key_up = nullptr;
    float2* tw14 = nullptr;
    float2* cls = nullptr;  // pass-A class tables (pa::CLS_TABLE_TOTAL entries)
    int cus = 256;  // compute units of the device (persistent kernels launch one workgroup each)
    std::map<uint32_t, Twiddles> tw;
    std::map<uint32_t, std::unique_ptr<Workspace>> ws;
    std::map<pa_seed_id, Seed> seeds;
    std::map<pa_ticket, Ticket> tickets;
    pa_seed_id next_seed = 1;
    pa_ticket next_ticket = 1;
    uint64_t used_bytes = 0;
    std::mutex mu;
    // stage timing
    bool timing = false;
    struct Pending { int stage; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    double stage_ms[16] = {};
    uint32_t stage_count[16] = {};
    // batched launches
    BatchSlot batch;                         // pa_batch_device / pa_run_batch
    struct Tuned { uint32_t grid_x, fwd_per; };
    std::map<std::string, Tuned> tuned;      // launch shape per (L, seed mode, count)
    struct Server* server = nullptr;         // pa_serve_*
    struct PinRange { uint64_t bytes; uint8_t* dev; };
    std::map<uintptr_t, PinRange> pinned;    // pa_host_alloc ranges: base -> size, mapped address
    uint8_t* split = nullptr;                // sub-block mode: tile inputs/outputs, band accumulator,
    uint64_t split_cap = 0;                  // host-API staging (grow-only)
    std::mutex pin_mu;
    uint8_t* noalloc = nullptr;              // this device's `scratch` address (pa_block_device_noalloc)
};

namespace {

int dev_alloc(pa_engine* e, void** p, size_t bytes) {
    if (e->cfg.hbm_budget_bytes && e->used_bytes + bytes > e->cfg.hbm_budget_bytes) return PA_ERR_NOMEM;
    hipError_t r = hipMalloc(p, bytes);
    if (r != hipSuccess) {
        set_error(r, "hipMalloc", __FILE__, __LINE__);
        (void)hipGetLastError();
        return PA_ERR_NOMEM;
    }
    e->used_bytes += bytes;
    return PA_OK;
}
void dev_free(pa_engine* e, void* p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    e->used_bytes -= bytes;
}

// Pinned host memory that kernels read and write directly over the bus (mapped, coherent): the
// staging copy kernel moves batch inputs and outputs through it without a DMA command per range.
hipError_t host_alloc_mapped(void** h, uint8_t** d, size_t bytes) {
    *h = nullptr;
    hipError_t r = hipHostMalloc(h, bytes, hipHostMallocMapped | hipHostMallocCoherent);
    if (r != hipSuccess) {
        *h = nullptr;
        return r;
    }
    r = hipHostGetDevicePointer(reinterpret_cast<void**>(d), *h, 0);
    if (r != hipSuccess) {
        (void)hipHostFree(*h);
        *h = nullptr;
    }
    return r;
}

// Grows a pinned, mapped host buffer to at least `need` bytes (contents not kept).
template <typename T>
hipError_t host_grow(T** h, uint8_t** d, uint64_t* cap, uint64_t need) {
    if (*cap >= need) return hipSuccess;
    if (*h) (void)hipHostFree(*h);
    *h = nullptr;
    *cap = 0;
    const uint64_t want = std::max<uint64_t>(need, need + need / 2);
    void* p = nullptr;
    hipError_t r = host_alloc_mapped(&p, d, want);
    if (r != hipSuccess) return r;
    *h = static_cast<T*>(p);
    *cap = want;
    return hipSuccess;
}

// Streams and events of a batch slot (own_stream: a launch stream of its own, for serving).
int slot_init(BatchSlot* sl, bool own_stream) {
    if (own_stream) PA_HIP(hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking));
    for (int i = 0; i < BatchSlot::FORKS; ++i) {
        PA_HIP(hipStreamCreateWithFlags(&sl->fork[i], hipStreamNonBlocking));
        PA_HIP(hipEventCreateWithFlags(&sl->join_ev[i], hipEventDisableTiming));
    }
    PA_HIP(hipEventCreateWithFlags(&sl->fork_ev, hipEventDisableTiming));
    PA_HIP(hipEventCreateWithFlags(&sl->htab_free, hipEventDisableTiming));
    return PA_OK;
}

void slot_release(pa_engine* e, BatchSlot* sl) {
    if (sl->stream) (void)hipStreamSynchronize(sl->stream);
    for (auto& kv : sl->graphs) (void)hipGraphExecDestroy(kv.second);
    sl->graphs.clear();
    for (int i = 0; i < BatchSlot::FORKS; ++i) {
        if (sl->fork[i]) {
            (void)hipStreamSynchronize(sl->fork[i]);
            (void)hipStreamDestroy(sl->fork[i]);
        }
        if (sl->join_ev[i]) (void)hipEventDestroy(sl->join_ev[i]);
    }
    if (sl->fork_ev) (void)hipEventDestroy(sl->fork_ev);
    if (sl->htab_free) (void)hipEventDestroy(sl->htab_free);
    if (sl->arena) dev_free(e, sl->arena, sl->arena_cap);
    if (sl->htab) (void)hipHostFree(sl->htab);
    if (sl->hdesc) (void)hipHostFree(sl->hdesc);
    if (sl->stream) (void)hipStreamDestroy(sl->stream);
    *sl = BatchSlot();
}

std::vector<float2> twiddles(uint64_t L, uint64_t count, uint64_t stride) {
    std::vector<float2> t(count);
    for (uint64_t i = 0; i < count; ++i) {
        const uint64_t x = (i * stride) % L;
        const double a = -2.0 * M_PI * (double)x / (double)L;
        t[i] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    return t;
}

// Pass A's class tables (layout in pa_kernels.hpp): sums of D-th roots of unity over the set bits
// of a byte, in double precision, rounded once.
std::vector<float2> class_tables() {
    std::vector<float2> t(pa::CLS_TABLE_TOTAL);
    for (int D = 1; D <= 16; D *= 2) {
        const int BW = pa::cls_bits(D), NB = pa::cls_tables(D);
        for (int c = 0; c < D; ++c)
            for (int q = 0; q < NB; ++q)
                for (int byte = 0; byte < (1 << BW); ++byte) {
                    double re = 0.0, im = 0.0;
                    for (int b = 0; b < BW; ++b)
                        if ((byte >> b) & 1) {
                            const double a = -2.0 * M_PI * (double)(((BW * q + b) * c) % D) / (double)D;
                            re += std::cos(a);
                            im += std::sin(a);
                        }
                    t[pa::cls_table_offset(D) + (c * NB + q) * (1 << BW) + byte] = make_float2((float)re, (float)im);
                }
    }
    return t;
}

uint64_t workspace_bytes(const Geom& g) {
    const uint64_t N = g.N;
    uint64_t b = (16384 + (N > 8192 ? g.L / 16384 : 1)) * sizeof(float2);
    if (g.N1 > 1) {
        b += N * sizeof(float2) + (N + 1) * sizeof(float2);
        b += 2 * (N / g.D) * 4 + pa::nib_bytes(g.N, g.W) + 192;
        if (pa::rt_tiled(g.lg1)) b += N * sizeof(float2);
    }
    b += 256 + ws_key_bytes(g.L) * 2 + ws_out_bytes(g.L);  // key/seed and output staging
    return b;
}

int get_twiddles(pa_engine* e, uint32_t L, Twiddles** out) {
    auto it = e->tw.find(L);
    if (it != e->tw.end()) {
        *out = &it->second;
        return PA_OK;
    }
    Twiddles t;
    const uint64_t hi_n = L > 16384 ? L / 16384 : 1;
    int rc = dev_alloc(e, (void**)&t.lo, 16384 * sizeof(float2));
    if (rc != PA_OK) return rc;
    rc = dev_alloc(e, (void**)&t.hi, hi_n * sizeof(float2));
    if (rc != PA_OK) {
        dev_free(e, t.lo, 16384 * sizeof(float2));
        return rc;
    }
    auto lo = twiddles(L, 16384, 1);
    auto hi = twiddles(L, hi_n, 16384);
    PA_HIP(hipMemcpy(t.lo, lo.data(), lo.size() * sizeof(float2), hipMemcpyHostToDevice));
    PA_HIP(hipMemcpy(t.hi, hi.data(), hi.size() * sizeof(float2), hipMemcpyHostToDevice));
    *out = &e->tw.emplace(L, t).first->second;
    return PA_OK;
}

int get_workspace(pa_engine* e, const Geom& g, Workspace** out) {
    auto it = e->ws.find(g.L);
    if (it != e->ws.end()) {
        *out = it->second.get();
        return PA_OK;
    }
    Twiddles* tw;
    int rc = get_twiddles(e, g.L, &tw);
    if (rc != PA_OK) return rc;
    auto w = std::make_unique<Workspace>();
    w->L = g.L;
    const size_t key_bytes = ws_key_bytes(g.L), out_bytes = ws_out_bytes(g.L);
    struct A { void** p; size_t b; };
    std::vector<A> allocs = {
        {(void**)&w->counts, 2 * pa::POP_SLOTS * sizeof(uint32_t)},
        {(void**)&w->d_key, key_bytes},
        {(void**)&w->d_seed, key_bytes},
        {(void**)&w->d_out, out_bytes},
    };
    if (g.N1 > 1) {
        allocs.push_back({(void**)&w->Yk, (size_t)g.N * sizeof(float2)});
        allocs.push_back({(void**)&w->Ys, (size_t)(g.N + 1) * sizeof(float2)});
        allocs.push_back({(void**)&w->tbk, (size_t)(g.N / g.D) * 4});
        allocs.push_back({(void**)&w->tbs, (size_t)(g.N / g.D) * 4});
        allocs.push_back({(void**)&w->nib, (size_t)pa::nib_bytes(g.N, g.W)});
        if (pa::rt_tiled(g.lg1)) allocs.push_back({(void**)&w->Rt, (size_t)g.N * sizeof(float2)});
    }
    for (auto& a : allocs) {
        rc = dev_alloc(e, a.p, a.b);
        if (rc != PA_OK) {
            for (auto& b : allocs)
                if (*b.p) dev_free(e, *b.p, b.b);
            return rc;
        }
        w->bytes += a.b;
    }
    PA_HIP(hipEventCreateWithFlags(&w->last_done, hipEventDisableTiming));
    *out = w.get();
    e->ws.emplace(g.L, std::move(w));
    return PA_OK;
}

// Twiddle tables of transform length L (get_twiddles must have created them).
pa::Tables tables(pa_engine* e, uint32_t L) {
    const Twiddles& t = e->tw.at(L);
    return pa::Tables{e->tw14, t.lo, t.hi, e->cls};
}

hipStream_t stream_of(pa_engine* e, void* s) { return s ? (hipStream_t)s : e->stream; }

enum Stage { ST_SMALL, ST_ZERO, ST_TRANSPOSE_KEY, ST_FWD_COLS_KEY, ST_TRANSPOSE_BOTH, ST_FWD_COLS_BOTH,
             ST_ROWS_SEEDPREP, ST_ROWS, ST_INV_COLS, ST_PACK, ST_TRANSPOSE_SEED, ST_FWD_COLS_SEED, ST_COUNT };
const char* const kStageNames[ST_COUNT] = {
    "k_small", "k_zero_u32", "k_transpose_bits(key)", "k_fwd_cols(key)", "k_transpose_bits(key+seed)",
    "k_fwd_cols(key+seed)", "k_rows(seedprep)", "k_rows", "k_inv_cols", "k_pack", "k_transpose_bits(seed)",
    "k_fwd_cols(seed)"};
static_assert(ST_COUNT <= 16, "stage tables hold 16 entries");

// A fresh seed still in host memory (pa_run_block): it is copied to dev on the upload stream
// after the key's copy (event key_up; both share the bus, so running them together would only
// delay the key) and while the key's first passes run; the seed's passes wait for it.
struct SeedUpload {
    const uint32_t* host;
    uint32_t* dev;
    size_t bytes;
    hipStream_t up;
    hipEvent_t key_up, done;
};

hipEvent_t pool_event(pa_engine* e) {
    if (!e->event_pool.empty()) {
        hipEvent_t ev = e->event_pool.back();
        e->event_pool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    return ev;
}

// Runs one launch, bracketed by events when stage timing is on.
template <typename F>
hipError_t timed(pa_engine* e, int stage, hipStream_t s, F&& launch) {
    if (!e->timing) return launch();
    hipEvent_t a = pool_event(e), b = pool_event(e);
    if (!a || !b) return hipErrorOutOfMemory;
    hipError_t r = hipEventRecord(a, s);
    if (r != hipSuccess) return r;
    r = launch();
    if (r != hipSuccess) return r;
    r = hipEventRecord(b, s);
    e->pending.push_back({stage, a, b});
    return r;
}

// The hot path on device buffers.  sspec/ws_dev != null selects the resident seed; su != null
// brings a fresh seed up from host memory during the key's passes (seed = su->dev).
int run_device_body(pa_engine* e, const Geom& g, Workspace* w, const uint32_t* key, const uint32_t* seed,
                    const float2* sspec, const uint32_t* ws_dev, uint32_t* out, hipStream_t s, float* margin,
                    const SeedUpload* su);

// Orders the workspace's uses across streams (not while s is being captured into a graph: a
// capture only records work, and the caller owns the ordering of its replays).
int run_device(pa_engine* e, const Geom& g, Workspace* w, const uint32_t* key, const uint32_t* seed,
               const float2* sspec, const uint32_t* ws_dev, uint32_t* out, hipStream_t s, float* margin,
               const SeedUpload* su = nullptr) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    PA_HIP(hipStreamIsCapturing(s, &cap));
    const bool guard = cap == hipStreamCaptureStatusNone;
    if (guard && w->last_stream && w->last_stream != s) PA_HIP(hipStreamWaitEvent(s, w->last_done, 0));
    const int rc = run_device_body(e, g, w, key, seed, sspec, ws_dev, out, s, margin, su);
    if (rc != PA_OK) return rc;
    if (guard) {
        PA_HIP(hipEventRecord(w->last_done, s));
        w->last_stream = s;
    }
    return PA_OK;
}

int run_device_body(pa_engine* e, const Geom& g, Workspace* w, const uint32_t* key, const uint32_t* seed,
                    const float2* sspec, const uint32_t* ws_dev, uint32_t* out, hipStream_t s, float* margin,
                    const SeedUpload* su) {
    const pa::Tables t = tables(e, g.L);
    if (g.N1 == 1) {
        PA_HIP(timed(e, ST_SMALL, s, [&] {
            return pa::launch_small(key, seed, sspec, ws_dev, out, g, t, margin, nullptr, nullptr, s);
        }));
        return PA_OK;
    }
    uint32_t* key_cnt = w->counts;
    uint32_t* seed_cnt = w->counts + pa::POP_SLOTS;
    if (!w->counts_zero)
        PA_HIP(timed(e, ST_ZERO, s, [&] { return pa::launch_zero_u32(w->counts, 2 * pa::POP_SLOTS, s); }));
    w->counts_zero = false;  // until this block's pack stage is enqueued
    if (!sspec && su) {  // key passes first, the seed's after its upload
        PA_HIP(timed(e, ST_TRANSPOSE_KEY, s, [&] { return pa::launch_transpose(key, w->tbk, key_cnt, g, 0, s); }));
        PA_HIP(timed(e, ST_FWD_COLS_KEY, s, [&] { return pa::launch_fwd_cols(w->tbk, w->Yk, g, t, s); }));
        PA_HIP(hipStreamWaitEvent(su->up, su->key_up, 0));
        PA_HIP(hipMemcpyAsync(su->dev, su->host, su->bytes, hipMemcpyHostToDevice, su->up));
        PA_HIP(hipEventRecord(su->done, su->up));
        PA_HIP(hipStreamWaitEvent(s, su->done, 0));
        PA_HIP(timed(e, ST_TRANSPOSE_SEED, s, [&] { return pa::launch_transpose(su->dev, w->tbs, seed_cnt, g, 1, s); }));
        PA_HIP(timed(e, ST_FWD_COLS_SEED, s, [&] { return pa::launch_fwd_cols(w->tbs, w->Ys, g, t, s); }));
        PA_HIP(timed(e, ST_ROWS, s, [&] { return pa::launch_rows(pa::ROWS_DYN, w->Yk, w->Ys, w->Rt, g, t, e->cus, s); }));
        ws_dev = seed_cnt;
    } else if (!sspec) {  // key and seed share each launch of the bit transpose and of pass A
        PA_HIP(timed(e, ST_TRANSPOSE_BOTH, s, [&] {
            return pa::launch_transpose(key, w->tbk, key_cnt, g, 0, s, seed, w->tbs);
        }));
        PA_HIP(timed(e, ST_FWD_COLS_BOTH, s, [&] { return pa::launch_fwd_cols(w->tbk, w->Yk, g, t, s, w->tbs, w->Ys); }));
        PA_HIP(timed(e, ST_ROWS, s, [&] { return pa::launch_rows(pa::ROWS_DYN, w->Yk, w->Ys, w->Rt, g, t, e->cus, s); }));
        ws_dev = seed_cnt;
    } else {
        float2* S = const_cast<float2*>(sspec);
        PA_HIP(timed(e, ST_TRANSPOSE_KEY, s, [&] { return pa::launch_transpose(key, w->tbk, key_cnt, g, 0, s); }));
        PA_HIP(timed(e, ST_FWD_COLS_KEY, s, [&] { return pa::launch_fwd_cols(w->tbk, w->Yk, g, t, s); }));
        PA_HIP(timed(e, ST_ROWS, s, [&] { return pa::launch_rows(pa::ROWS_STATIC, w->Yk, S, w->Rt, g, t, e->cus, s); }));
    }
    const float2* R = pa::rt_tiled(g.lg1) ? w->Rt : w->Yk;
    PA_HIP(timed(e, ST_INV_COLS, s, [&] { return pa::launch_inv_cols(R, w->nib, g, key_cnt, ws_dev, t, margin, s); }));
    PA_HIP(timed(e, ST_PACK, s, [&] { return pa::launch_pack(w->nib, key, out, g, s, w->counts); }));
    w->counts_zero = true;  // the pack stage clears the partials for the next block on the stream
    return PA_OK;
}

int prepare(pa_engine* e, uint32_t n, uint32_t m, uint32_t flags, Geom* g, Workspace** w) {
    int rc = make_geom(n, m, flags, e->cfg.out_reverse_endian, e->cfg.max_n_log2, g);
    if (rc != PA_OK) return rc;
    if (!block_extents(*g).ok()) return PA_ERR_UNSUPPORTED;  // a kernel would index past a buffer
    return get_workspace(e, *g, w);
}

}  // namespace

extern "C" {

int pa_config_default(pa_config* cfg) {
    if (!cfg) return PA_ERR_ARG;
    memset(cfg, 0, sizeof(*cfg));
    cfg->device = 0;
    cfg->max_n_log2 = 27;
    cfg->reduction_exp = 11;
    cfg->pre_mul_reduction_exp = 5;
    cfg->hbm_budget_bytes = 64ull << 30;
    cfg->out_reverse_endian = 1;
    return PA_OK;
}

int pa_create(const pa_config* cfg, pa_engine** out) {
    if (!out) return PA_ERR_ARG;
    auto e = std::make_unique<pa_engine>();
    if (cfg) e->cfg = *cfg; else pa_config_default(&e->cfg);
    if (e->cfg.max_n_log2 == 0) e->cfg.max_n_log2 = 27;
    PA_HIP(hipSetDevice(e->cfg.device));
    PA_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    PA_HIP(hipStreamCreateWithFlags(&e->upload, hipStreamNonBlocking));
    PA_HIP(hipEventCreateWithFlags(&e->upload_done, hipEventDisableTiming));
    PA_HIP(hipEventCreateWithFlags(&e->key_up, hipEventDisableTiming));
    PA_HIP(hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, e->cfg.device));
    int rc = slot_init(&e->batch, false);
    if (rc != PA_OK) return rc;
    rc = dev_alloc(e.get(), (void**)&e->tw14, 16384 * sizeof(float2));
    if (rc != PA_OK) return rc;
    auto tw = twiddles(16384, 16384, 1);
    PA_HIP(hipMemcpy(e->tw14, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    rc = dev_alloc(e.get(), (void**)&e->cls, pa::CLS_TABLE_TOTAL * sizeof(float2));
    if (rc != PA_OK) return rc;
    auto cls = class_tables();
    PA_HIP(hipMemcpy(e->cls, cls.data(), cls.size() * sizeof(float2), hipMemcpyHostToDevice));
    // pa_block_device_noalloc: the scratch array's address on this device, and the twiddle tables
    // of every four-step length it serves, so that its calls allocate nothing
    {
        void* sp = nullptr;
        PA_HIP(scratch_address(e->cfg.device, &sp));
        e->noalloc = static_cast<uint8_t*>(sp);
        for (uint32_t lg = pa::SMALL_MAX_LOG_N + 2; lg <= std::min(NOALLOC_MAX_L_LOG, e->cfg.max_n_log2); ++lg) {
            Twiddles* tw;
            rc = get_twiddles(e.get(), 1u << lg, &tw);
            if (rc != PA_OK) return rc;
        }
    }
    *out = e.release();
    return PA_OK;
}

int pa_destroy(pa_engine* e) {
    if (!e) return PA_ERR_ARG;
    (void)hipSetDevice(e->cfg.device);
    // The serving scheduler launches kernels that read resident seeds, workspaces and twiddle
    // tables: drain its queued and in-flight blocks and stop it before anything is freed.
    if (e->server) pa_serve_stop(e);
    std::unique_lock<std::mutex> lk(e->mu);
    (void)hipStreamSynchronize(e->stream);
    for (auto& kv : e->tickets) {
        Ticket& t = kv.second;
        if (t.done) (void)hipEventDestroy(t.done);
        (void)hipFree(t.dev_out);
        (void)hipFree(t.dev_key);
        (void)hipFree(t.dev_seed);
    }
    for (auto& kv : e->seeds) {
        (void)hipFree(kv.second.spec);
        (void)hipFree(kv.second.ws);
    }
    for (auto& kv : e->ws) {
        Workspace* w = kv.second.get();
        void* ps[] = {w->Yk, w->Ys, w->Rt, w->tbk, w->tbs, w->nib, w->counts, w->d_key, w->d_seed, w->d_out};
        for (void* p : ps) (void)hipFree(p);
        if (w->last_done) (void)hipEventDestroy(w->last_done);
    }
    for (auto& kv : e->tw) {
        (void)hipFree(kv.second.lo);
        (void)hipFree(kv.second.hi);
    }
    for (auto& p : e->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t ev : e->event_pool) (void)hipEventDestroy(ev);
    for (auto& kv : e->pinned) (void)hipHostFree(reinterpret_cast<void*>(kv.first));
    slot_release(e, &e->batch);
    if (e->split) (void)hipFree(e->split);
    (void)hipFree(e->tw14);
    (void)hipFree(e->cls);
    (void)hipStreamDestroy(e->stream);
    (void)hipStreamSynchronize(e->upload);
    (void)hipStreamDestroy(e->upload);
    (void)hipEventDestroy(e->upload_done);
    (void)hipEventDestroy(e->key_up);
    lk.unlock();  // the mutex is a member of *e
    delete e;
    return PA_OK;
}

uint64_t pa_workspace_bytes(uint32_t n, uint32_t m, uint32_t flags) {
    Geom g;
    if (make_geom(n, m, flags, 1, 31, &g) != PA_OK) return 0;
    return workspace_bytes(g);
}

int pa_block_extents(uint32_t n, uint32_t m, uint32_t flags, uint32_t* transform_len,
                     uint64_t touched[PA_BUF_COUNT], uint64_t alloc[PA_BUF_COUNT]) {
    Geom g;
    const int rc = make_geom(n, m, flags, 1, pa::MAX_L_LOG, &g);
    if (rc != PA_OK) return rc;
    const Extents x = block_extents(g);
    if (transform_len) *transform_len = g.L;
    for (int i = 0; i < PA_BUF_COUNT; ++i) {
        if (touched) touched[i] = x.touched[i];
        if (alloc) alloc[i] = x.alloc[i];
    }
    return PA_OK;
}

int pa_seed_put(pa_engine* e, const uint32_t* seed, uint32_t n, uint32_t m, int32_t reuse, pa_seed_id* id) {
    if (!e || !seed || !id) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    Geom g;
    Workspace* w;
    int rc = prepare(e, n, m, PA_DEFAULT_FLAGS, &g, &w);
    if (rc != PA_OK) return rc;
    Seed sd;
    sd.g = g;
    sd.reuse = reuse;
    sd.bytes = (uint64_t)(g.N + 1) * sizeof(float2) + pa::POP_SLOTS * sizeof(uint32_t);
    rc = dev_alloc(e, (void**)&sd.spec, (size_t)(g.N + 1) * sizeof(float2));
    if (rc != PA_OK) return rc;
    rc = dev_alloc(e, (void**)&sd.ws, pa::POP_SLOTS * sizeof(uint32_t));
    if (rc != PA_OK) {
        dev_free(e, sd.spec, (size_t)(g.N + 1) * sizeof(float2));
        return rc;
    }
    hipStream_t s = e->stream;
    PA_HIP(hipMemcpyAsync(w->d_seed, seed, (size_t)(n / 32) * 4, hipMemcpyHostToDevice, s));
    const pa::Tables t = tables(e, g.L);
    if (g.N1 == 1) {
        PA_HIP(pa::launch_small(nullptr, w->d_seed, nullptr, nullptr, nullptr, g, t, nullptr, sd.spec, sd.ws, s));
    } else {
        PA_HIP(pa::launch_zero_u32(sd.ws, pa::POP_SLOTS, s));
        PA_HIP(pa::launch_transpose(w->d_seed, w->tbs, sd.ws, g, 1, s));
        PA_HIP(pa::launch_fwd_cols(w->tbs, sd.spec, g, t, s));
        PA_HIP(pa::launch_rows(pa::ROWS_SEEDPREP, nullptr, sd.spec, nullptr, g, t, e->cus, s));
    }
    PA_HIP(hipStreamSynchronize(s));
    *id = e->next_seed++;
    e->seeds[*id] = sd;
    return PA_OK;
}

int pa_seed_drop(pa_engine* e, pa_seed_id id) {
    if (!e) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    auto it = e->seeds.find(id);
    if (it == e->seeds.end()) return PA_ERR_NOT_FOUND;
    (void)hipStreamSynchronize(e->stream);
    dev_free(e, it->second.spec, (size_t)(it->second.g.N + 1) * sizeof(float2));
    dev_free(e, it->second.ws, pa::POP_SLOTS * sizeof(uint32_t));
    e->seeds.erase(it);
    return PA_OK;
}

static int seed_lookup(pa_engine* e, pa_seed_id id, const Geom& g, const Seed** out) {
    auto it = e->seeds.find(id);
    if (it == e->seeds.end()) return PA_ERR_NOT_FOUND;
    const Geom& sg = it->second.g;
    if (sg.n != g.n || sg.L != g.L || sg.seed_rot != g.seed_rot) return PA_ERR_ARG;
    *out = &it->second;
    return PA_OK;
}

int pa_block_device(pa_engine* e, const uint32_t* d_key, const uint32_t* d_seed, uint32_t n, uint32_t m,
                    uint32_t flags, uint8_t* d_out, void* stream, float* d_margin) {
    if (!e || !d_key || !d_seed || !d_out) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    Geom g;
    Workspace* w;
    int rc = prepare(e, n, m, flags, &g, &w);
    if (rc != PA_OK) return rc;
    return run_device(e, g, w, d_key, d_seed, nullptr, nullptr, (uint32_t*)d_out, stream_of(e, stream), d_margin);
}

int pa_block_device_seeded(pa_engine* e, const uint32_t* d_key, pa_seed_id id, uint32_t n, uint32_t m,
                           uint32_t flags, uint8_t* d_out, void* stream, float* d_margin) {
    if (!e || !d_key || !d_out) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    Geom g;
    Workspace* w;
    int rc = prepare(e, n, m, flags, &g, &w);
    if (rc != PA_OK) return rc;
    const Seed* sd;
    rc = seed_lookup(e, id, g, &sd);
    if (rc != PA_OK) return rc;
    return run_device(e, g, w, d_key, nullptr, sd->spec, sd->ws, (uint32_t*)d_out, stream_of(e, stream), d_margin);
}

int pa_block_device_noalloc(pa_engine* e, const uint32_t* d_key, const uint32_t* d_seed, uint32_t n, uint32_t m,
                            uint32_t flags, uint8_t* d_out, void* stream, float* d_margin) {
    if (!e || !d_key || !d_seed || !d_out) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    Geom g;
    int rc = make_geom(n, m, flags, e->cfg.out_reverse_endian, e->cfg.max_n_log2, &g);
    if (rc != PA_OK) return rc;
    if (g.L > (1u << NOALLOC_MAX_L_LOG)) {
        snprintf(g_size_error, sizeof(g_size_error),
                 "pa_block_device_noalloc serves transform lengths up to 2^%u bits (n <= 2^%u); n = %u needs 2^%u",
                 NOALLOC_MAX_L_LOG, NOALLOC_MAX_L_LOG, n, log2u(g.L));
        return PA_ERR_SIZE;
    }
    if (!block_extents(g).ok()) return PA_ERR_UNSUPPORTED;
    if (!e->noalloc || e->tw.find(g.L) == e->tw.end()) return PA_ERR_UNSUPPORTED;
    // the workspace of this block, carved from the scratch array (no allocation); the popcount
    // partials are cleared by a kernel at the start of the block (counts_zero = false)
    Workspace w;
    w.L = g.L;
    if (g.N1 > 1) {
        uint64_t off = 0;
        auto take = [&](uint64_t bytes) {
            uint8_t* q = e->noalloc + off;
            off += (bytes + 255u) & ~uint64_t(255);
            return q;
        };
        w.Yk = reinterpret_cast<float2*>(take((uint64_t)g.N * 8));
        w.Ys = reinterpret_cast<float2*>(take((uint64_t)(g.N + 1) * 8));
        w.tbk = reinterpret_cast<uint32_t*>(take((uint64_t)(g.N / g.D) * 4));
        w.tbs = reinterpret_cast<uint32_t*>(take((uint64_t)(g.N / g.D) * 4));
        w.nib = take(pa::nib_bytes(g.N, g.W));
        w.counts = reinterpret_cast<uint32_t*>(take(2 * pa::POP_SLOTS * 4));
        if (pa::rt_tiled(g.lg1)) w.Rt = reinterpret_cast<float2*>(take((uint64_t)g.N * 8));
        if (off > NOALLOC_SCRATCH_BYTES) return PA_ERR_UNSUPPORTED;
    }
    const bool timing = e->timing;  // stage timing creates events: not on this path
    e->timing = false;
    rc = run_device_body(e, g, &w, d_key, d_seed, nullptr, nullptr, reinterpret_cast<uint32_t*>(d_out),
                         stream_of(e, stream), d_margin, nullptr);
    e->timing = timing;
    return rc;
}

static int run_host(pa_engine* e, const uint32_t* key, const uint32_t* seed, pa_seed_id id, uint32_t n,
                    uint32_t m, uint32_t flags, uint8_t* out) {
    if (!e || !key || !out || (!seed && !id)) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    Geom g;
    Workspace* w;
    int rc = prepare(e, n, m, flags, &g, &w);
    if (rc != PA_OK) return rc;
    const Seed* sd = nullptr;
    if (!seed) {
        rc = seed_lookup(e, id, g, &sd);
        if (rc != PA_OK) return rc;
    }
    hipStream_t s = e->stream;
    const uint32_t kw = key_words_for(g);
    // zero the words after the key so key-rest reads past a NoXOR key stay defined
    PA_HIP(hipMemsetAsync(w->d_key + kw, 0, (size_t)(g.n / 32 + 2 - kw) * 4, s));
    PA_HIP(hipMemcpyAsync(w->d_key, key, (size_t)kw * 4, hipMemcpyHostToDevice, s));
    // a fresh seed of a multi-pass block goes up while the key's first passes run (the previous
    // call's kernels, the only other readers of d_seed, finished before it returned)
    const SeedUpload su{seed, w->d_seed, (size_t)(n / 32) * 4, e->upload, e->key_up, e->upload_done};
    const bool overlap = seed && g.N1 > 1;
    if (overlap) PA_HIP(hipEventRecord(e->key_up, s));
    if (seed && !overlap) PA_HIP(hipMemcpyAsync(w->d_seed, seed, (size_t)(n / 32) * 4, hipMemcpyHostToDevice, s));
    rc = run_device(e, g, w, w->d_key, seed ? w->d_seed : nullptr, sd ? sd->spec : nullptr,
                    sd ? sd->ws : nullptr, w->d_out, s, nullptr, overlap ? &su : nullptr);
    if (rc != PA_OK) return rc;
    PA_HIP(hipMemcpyAsync(out, w->d_out, out_bytes_of(g), hipMemcpyDeviceToHost, s));
    PA_HIP(hipStreamSynchronize(s));
    return PA_OK;
}

int pa_run_block(pa_engine* e, const uint32_t* key, const uint32_t* seed, uint32_t n, uint32_t m,
                 uint32_t flags, uint8_t* out) {
    if (!seed) return PA_ERR_ARG;
    return run_host(e, key, seed, 0, n, m, flags, out);
}

int pa_run_block_seeded(pa_engine* e, const uint32_t* key, pa_seed_id id, uint32_t n, uint32_t m,
                        uint32_t flags, uint8_t* out) {
    if (!id) return PA_ERR_ARG;
    return run_host(e, key, nullptr, id, n, m, flags, out);
}

int pa_submit(pa_engine* e, const uint32_t* key, const uint32_t* seed, pa_seed_id id, uint32_t n, uint32_t m,
              uint32_t flags, uint8_t* out, pa_ticket* ticket) {
    if (!e || !key || !out || !ticket || (!seed && !id)) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    Geom g;
    Workspace* w;
    int rc = prepare(e, n, m, flags, &g, &w);
    if (rc != PA_OK) return rc;
    const Seed* sd = nullptr;
    if (!seed) {
        rc = seed_lookup(e, id, g, &sd);
        if (rc != PA_OK) return rc;
    }
    hipStream_t s = e->stream;
    Ticket t;
    t.host_out = out;
    t.out_bytes = out_bytes_of(g);
    const size_t kbytes = stage_key_bytes(g);
    PA_HIP(hipMalloc(&t.dev_key, kbytes));
    PA_HIP(hipMalloc(&t.dev_out, stage_out_bytes(g)));
    if (seed) PA_HIP(hipMalloc(&t.dev_seed, stage_seed_bytes(g)));
    PA_HIP(hipMemsetAsync(t.dev_key, 0, kbytes, s));
    PA_HIP(hipMemcpyAsync(t.dev_key, key, (size_t)key_words_for(g) * 4, hipMemcpyHostToDevice, s));
    if (seed) PA_HIP(hipMemcpyAsync(t.dev_seed, seed, stage_seed_bytes(g), hipMemcpyHostToDevice, s));
    rc = run_device(e, g, w, t.dev_key, t.dev_seed, sd ? sd->spec : nullptr, sd ? sd->ws : nullptr,
                    t.dev_out, s, nullptr);
    if (rc != PA_OK) return rc;
    PA_HIP(hipMemcpyAsync(out, t.dev_out, t.out_bytes, hipMemcpyDeviceToHost, s));
    PA_HIP(hipEventCreateWithFlags(&t.done, hipEventDisableTiming));
    PA_HIP(hipEventRecord(t.done, s));
    *ticket = e->next_ticket++;
    e->tickets[*ticket] = t;
    return PA_OK;
}

int pa_wait(pa_engine* e, pa_ticket ticket, int64_t timeout_ns) {
    if (!e) return PA_ERR_ARG;
    hipEvent_t ev;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        auto it = e->tickets.find(ticket);
        if (it == e->tickets.end()) return PA_ERR_NOT_FOUND;
        if (it->second.waiting) return PA_ERR_ARG;  // one waiter per ticket: the winner destroys the event
        it->second.waiting = true;
        ev = it->second.done;
    }
    auto give_up = [&](int st) {
        std::lock_guard<std::mutex> lk(e->mu);
        auto it = e->tickets.find(ticket);
        if (it != e->tickets.end()) it->second.waiting = false;
        return st;
    };
    const auto t0 = std::chrono::steady_clock::now();
    while (true) {
        hipError_t q = hipEventQuery(ev);
        if (q == hipSuccess) break;
        if (q != hipErrorNotReady) {
            set_error(q, "hipEventQuery", __FILE__, __LINE__);
            return give_up(PA_ERR_DEVICE);
        }
        if (timeout_ns >= 0) {
            const auto dt = std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0);
            if (dt.count() > timeout_ns) return give_up(PA_ERR_TIMEOUT);
        }
        std::this_thread::yield();
    }
    std::lock_guard<std::mutex> lk(e->mu);
    auto it = e->tickets.find(ticket);
    if (it == e->tickets.end()) return PA_ERR_NOT_FOUND;  // another waiter retired it meanwhile
    Ticket t = it->second;
    e->tickets.erase(it);
    (void)hipEventDestroy(t.done);
    (void)hipFree(t.dev_out);
    (void)hipFree(t.dev_key);
    (void)hipFree(t.dev_seed);
    return PA_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batched launches ----
namespace {

uint64_t align256(uint64_t x) { return (x + 255u) & ~uint64_t(255); }

struct BatchItem {
    Geom g{};
    const Seed* sd = nullptr;  // resident seed (static mode)
    bool dyn = true;
};

struct BatchGroup {
    uint32_t L = 0;
    bool dyn = true, small = false;
    std::vector<uint32_t> idx;  // blocks of the group (caller order)
    uint32_t tab0 = 0;          // first table row of the group
    uint32_t tiles_y = 1;       // pack grid.y: the largest window of the group
    uint32_t grid_x = 0;        // row-pass workgroups per block
    uint32_t fwd_per = 0;       // pass A walk length (0: the launcher's default)
    double cost = 0.0;
};

struct BatchPlan {
    std::vector<BatchItem> items;
    std::vector<BatchGroup> groups;
    uint64_t table_bytes = 0, scratch_bytes = 0;
};

// Device scratch of one block of a batch (the four-step path's intermediate buffers).
uint64_t batch_scratch_bytes(const Geom& g, bool dyn) {
    if (g.N1 == 1) return 0;
    uint64_t b = align256((uint64_t)g.N * 8) + align256((uint64_t)(g.N / g.D) * 4);  // Yk, tbk
    if (dyn) b += align256((uint64_t)(g.N + 1) * 8) + align256((uint64_t)(g.N / g.D) * 4);  // Ys, tbs
    b += align256(pa::nib_bytes(g.N, g.W)) + align256(2 * pa::POP_SLOTS * 4);  // nib, counts
    if (pa::rt_tiled(g.lg1)) b += align256((uint64_t)g.N * 8);     // R' tiles
    return b;
}

int batch_plan(pa_engine* e, const pa_block* blocks, uint32_t count, BatchPlan* p) {
    if (!blocks || count == 0 || count > 65535u) return PA_ERR_ARG;
    p->items.resize(count);
    std::map<std::pair<uint32_t, bool>, uint32_t> gid;
    for (uint32_t i = 0; i < count; ++i) {
        const pa_block& b = blocks[i];
        if (!b.key || !b.out || (!b.seed && !b.seed_id) || b.reserved) return PA_ERR_ARG;
        BatchItem& it = p->items[i];
        int rc = make_geom(b.n, b.m, b.flags, e->cfg.out_reverse_endian, e->cfg.max_n_log2, &it.g);
        if (rc != PA_OK) return rc;
        it.dyn = b.seed != nullptr;
        if (!it.dyn) {
            rc = seed_lookup(e, b.seed_id, it.g, &it.sd);
            if (rc != PA_OK) return rc;
        }
        Twiddles* tw;
        rc = get_twiddles(e, it.g.L, &tw);
        if (rc != PA_OK) return rc;
        const auto key = std::make_pair(it.g.L, it.dyn);
        auto f = gid.find(key);
        if (f == gid.end()) {
            f = gid.emplace(key, (uint32_t)p->groups.size()).first;
            BatchGroup gr;
            gr.L = it.g.L;
            gr.dyn = it.dyn;
            gr.small = it.g.N1 == 1;
            p->groups.push_back(gr);
        }
        BatchGroup& gr = p->groups[f->second];
        gr.idx.push_back(i);
        gr.cost += (double)it.g.L * (it.dyn ? 1.5 : 1.0);
        if (!gr.small) gr.tiles_y = std::max(gr.tiles_y, pa::pack_tiles(it.g));
        p->scratch_bytes += batch_scratch_bytes(it.g, it.dyn);
    }
    uint32_t row = 0;
    for (auto& gr : p->groups) {
        gr.tab0 = row;
        row += (uint32_t)gr.idx.size();
    }
    p->table_bytes = align256((uint64_t)count * sizeof(pa::BlockArgs));
    return PA_OK;
}

// Grows the batch arena to at least `bytes` (drains the device first; cached graphs refer to the
// old arena and are dropped).
int ensure_arena(pa_engine* e, BatchSlot& sl, uint64_t bytes, uint64_t limit = 0) {
    if (bytes <= sl.arena_cap) return PA_OK;
    if (limit && bytes > limit) return PA_ERR_NOMEM;
    if (sl.stream) PA_HIP(hipStreamSynchronize(sl.stream));
    else PA_HIP(hipDeviceSynchronize());
    for (auto& kv : sl.graphs) (void)hipGraphExecDestroy(kv.second);
    sl.graphs.clear();
    if (sl.arena) dev_free(e, sl.arena, sl.arena_cap);
    sl.arena = nullptr;
    sl.arena_cap = 0;
    // half again as much, so that a stream of growing batches reallocates only a few times
    uint64_t cap = align256(bytes + bytes / 2);
    if (limit && cap > limit) cap = align256(bytes);
    int rc = dev_alloc(e, (void**)&sl.arena, cap);
    if (rc == PA_ERR_NOMEM && cap > align256(bytes)) {
        cap = align256(bytes);
        rc = dev_alloc(e, (void**)&sl.arena, cap);
    }
    if (rc != PA_OK) return rc;
    sl.arena_cap = cap;
    return PA_OK;
}

// All launches of one group on stream s.
int batch_group_launch(pa_engine* e, const BatchPlan& p, const BatchGroup& gr, const pa::BlockArgs* tab,
                       uint32_t grid_x, uint32_t fwd_per, hipStream_t s) {
    const pa::BlockArgs* bt = tab + gr.tab0;
    const uint32_t cnt = (uint32_t)gr.idx.size();
    const Geom& g0 = p.items[gr.idx[0]].g;
    const pa::Tables t = tables(e, gr.L);
    if (gr.small) {
        PA_HIP(pa::launch_batch_small(bt, cnt, g0, t, s));
        return PA_OK;
    }
    PA_HIP(pa::launch_batch_zero(bt, cnt, s));
    PA_HIP(pa::launch_batch_transpose(bt, cnt, g0, 0, s));
    PA_HIP(pa::launch_batch_fwd_cols(bt, cnt, g0, t, 0, fwd_per, s));
    if (gr.dyn) {
        PA_HIP(pa::launch_batch_transpose(bt, cnt, g0, 1, s));
        PA_HIP(pa::launch_batch_fwd_cols(bt, cnt, g0, t, 1, fwd_per, s));
    }
    PA_HIP(pa::launch_batch_rows(gr.dyn ? pa::ROWS_DYN : pa::ROWS_STATIC, bt, cnt, g0, t, grid_x, s));
    PA_HIP(pa::launch_batch_inv_cols(bt, cnt, g0, t, s));
    PA_HIP(pa::launch_batch_pack(bt, cnt, g0, gr.tiles_y, s));
    return PA_OK;
}

// Groups on the fork streams (largest first onto the least-loaded stream), joined back into s.
int batch_launch_all(pa_engine* e, BatchSlot& sl, const BatchPlan& p, const pa::BlockArgs* tab, hipStream_t s) {
    const size_t ng = p.groups.size();
    if (ng == 1) return batch_group_launch(e, p, p.groups[0], tab, p.groups[0].grid_x, p.groups[0].fwd_per, s);
    std::vector<size_t> order(ng);
    for (size_t i = 0; i < ng; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return p.groups[a].cost > p.groups[b].cost; });
    double load[BatchSlot::FORKS] = {};
    bool used[BatchSlot::FORKS] = {};
    PA_HIP(hipEventRecord(sl.fork_ev, s));
    for (size_t gi : order) {
        int k = 0;
        for (int j = 1; j < BatchSlot::FORKS; ++j)
            if (load[j] < load[k]) k = j;
        if (!used[k]) {
            PA_HIP(hipStreamWaitEvent(sl.fork[k], sl.fork_ev, 0));
            used[k] = true;
        }
        load[k] += p.groups[gi].cost;
        int rc = batch_group_launch(e, p, p.groups[gi], tab, p.groups[gi].grid_x, p.groups[gi].fwd_per, sl.fork[k]);
        if (rc != PA_OK) return rc;
    }
    for (int k = 0; k < BatchSlot::FORKS; ++k) {
        if (!used[k]) continue;
        PA_HIP(hipEventRecord(sl.join_ev[k], sl.fork[k]));
        PA_HIP(hipStreamWaitEvent(s, sl.join_ev[k], 0));
    }
    return PA_OK;
}

std::string tune_key(const BatchGroup& gr) {
    return std::to_string(gr.L) + (gr.dyn ? "d" : "s") + std::to_string(gr.idx.size());
}

// The autotuner: times a group's whole launch sequence (the outputs are rewritten by the real run
// that follows) over its launch-shape knobs and records the fastest shape per (L, seed mode,
// count).  Knobs, tuned one after the other: the pass A walk length (column groups per persistent
// workgroup, for N1 = 2^9 .. 2^13), then the row-pass grid (workgroups per block).
int batch_tune_group(pa_engine* e, const BatchPlan& p, BatchGroup& gr, const pa::BlockArgs* tab, hipStream_t s) {
    const Geom& g0 = p.items[gr.idx[0]].g;
    const uint32_t items = g0.N1 / 2u;
    const uint32_t cnt = (uint32_t)gr.idx.size();
    const uint32_t base = std::max<uint32_t>(1u, ((uint32_t)e->cus + cnt - 1u) / cnt);
    hipEvent_t a, b;
    PA_HIP(hipEventCreate(&a));
    PA_HIP(hipEventCreate(&b));
    auto time_shape = [&](uint32_t x, uint32_t per, float* ms) -> int {
        for (int rep = 0; rep < 2; ++rep) {  // the first run warms up
            PA_HIP(hipEventRecord(a, s));
            int rc = batch_group_launch(e, p, gr, tab, x, per, s);
            if (rc != PA_OK) return rc;
            PA_HIP(hipEventRecord(b, s));
            PA_HIP(hipEventSynchronize(b));
            PA_HIP(hipEventElapsedTime(ms, a, b));
        }
        return PA_OK;
    };
    int rc = PA_OK;
    uint32_t best_x = std::min(base, items), best_per = 0;
    if (g0.lg1 >= 9 && g0.lg1 <= 13) {
        // walk lengths that leave at least one workgroup per compute unit (and the default)
        const uint32_t wgs = (g0.N2 / 16u) * g0.D * cnt;
        float best = 1e30f;
        for (uint32_t per = 1; per <= pa::FWD_PER_MAX && rc == PA_OK; per *= 2) {
            if (per > 1u && wgs / per < (uint32_t)e->cus) break;
            float ms = 0.f;
            rc = time_shape(best_x, per, &ms);
            if (rc == PA_OK && ms < best) {
                best = ms;
                best_per = per;
            }
        }
    }
    float best = 1e30f;
    for (uint32_t x = 1; x <= items && rc == PA_OK; x *= 2) {
        if (x * 4u < base || x > base * 8u) continue;
        float ms = 0.f;
        rc = time_shape(x, best_per, &ms);
        if (rc == PA_OK && ms < best) {
            best = ms;
            best_x = x;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc != PA_OK) return rc;
    e->tuned[tune_key(gr)] = {best_x, best_per};
    gr.grid_x = best_x;
    gr.fwd_per = best_per;
    return PA_OK;
}

// Fills and uploads the block table, then launches (tuning / graph replay as asked).
// `dev` holds the device pointers of key, seed, out and margin of every block.  `in` ranges are
// copied with the table before the transforms, `out` ranges after them, one launch each.
int batch_execute(pa_engine* e, BatchSlot& sl, BatchPlan& p, const pa_block* dev, uint32_t bflags, hipStream_t s,
                  const std::vector<pa::CopyDesc>* in = nullptr, const std::vector<pa::CopyDesc>* out = nullptr) {
    const uint32_t count = (uint32_t)p.items.size();
    pa::BlockArgs* dtab = reinterpret_cast<pa::BlockArgs*>(sl.arena);
    uint8_t* scratch = sl.arena + p.table_bytes;
    if (sl.htab_pending) {
        PA_HIP(hipEventSynchronize(sl.htab_free));
        sl.htab_pending = false;
    }
    {
        uint64_t cap = (uint64_t)sl.htab_cap * sizeof(pa::BlockArgs);
        PA_HIP(host_grow(&sl.htab, &sl.htab_dev, &cap, (uint64_t)count * sizeof(pa::BlockArgs)));
        sl.htab_cap = (uint32_t)(cap / sizeof(pa::BlockArgs));
        const size_t nd = 1 + (in ? in->size() : 0) + (out ? out->size() : 0);
        uint64_t dcap = (uint64_t)sl.hdesc_cap * sizeof(pa::CopyDesc);
        uint8_t* ddev = nullptr;
        const bool grow = dcap < nd * sizeof(pa::CopyDesc);
        PA_HIP(host_grow(&sl.hdesc, &ddev, &dcap, nd * sizeof(pa::CopyDesc)));
        if (grow) sl.hdesc_dev = reinterpret_cast<const pa::CopyDesc*>(ddev);
        sl.hdesc_cap = (uint32_t)(dcap / sizeof(pa::CopyDesc));
    }
    uint64_t off = 0;
    for (auto& gr : p.groups) {
        for (size_t j = 0; j < gr.idx.size(); ++j) {
            const uint32_t i = gr.idx[j];
            const BatchItem& it = p.items[i];
            pa::BlockArgs a{};
            a.g = it.g;
            a.key = dev[i].key;
            a.seed = it.dyn ? dev[i].seed : nullptr;
            a.sspec = it.sd ? it.sd->spec : nullptr;
            a.ws = it.sd ? it.sd->ws : nullptr;
            a.out = reinterpret_cast<uint32_t*>(dev[i].out);
            a.margin = dev[i].margin;
            if (!gr.small) {
                const Geom& g = it.g;
                auto take = [&](uint64_t bytes) {
                    uint8_t* q = scratch + off;
                    off += align256(bytes);
                    return q;
                };
                a.Yk = reinterpret_cast<float2*>(take((uint64_t)g.N * 8));
                a.tbk = reinterpret_cast<uint32_t*>(take((uint64_t)(g.N / g.D) * 4));
                if (it.dyn) {
                    a.Ys = reinterpret_cast<float2*>(take((uint64_t)(g.N + 1) * 8));
                    a.tbs = reinterpret_cast<uint32_t*>(take((uint64_t)(g.N / g.D) * 4));
                }
                a.nib = take(pa::nib_bytes(g.N, g.W));
                a.counts = reinterpret_cast<uint32_t*>(take(2 * pa::POP_SLOTS * 4));
                if (pa::rt_tiled(g.lg1)) a.Rt = reinterpret_cast<float2*>(take((uint64_t)g.N * 8));
            }
            sl.htab[gr.tab0 + j] = a;
        }
        if (!gr.small) {
            auto t = e->tuned.find(tune_key(gr));
            const uint32_t items = p.items[gr.idx[0]].g.N1 / 2u;
            const uint32_t cnt = (uint32_t)gr.idx.size();
            gr.grid_x = t != e->tuned.end() ? t->second.grid_x
                                            : std::min(items, std::max<uint32_t>(1u, ((uint32_t)e->cus + cnt - 1u) / cnt));
            gr.fwd_per = t != e->tuned.end() ? t->second.fwd_per : 0u;
        }
    }
    // the table and the input ranges in one copy launch (read from mapped host memory)
    uint32_t nd = 0;
    uint64_t longest = (uint64_t)count * sizeof(pa::BlockArgs);
    sl.hdesc[nd++] = {sl.htab_dev, reinterpret_cast<uint8_t*>(dtab), longest};
    if (in)
        for (const pa::CopyDesc& c : *in) {
            sl.hdesc[nd++] = c;
            longest = std::max(longest, c.bytes);
        }
    PA_HIP(pa::launch_copies(sl.hdesc_dev, nd, longest, s));
    sl.htab_pending = true;  // from here on htab_free must be recorded before returning
    auto finish = [&](int rc) -> int {
        if (rc == PA_OK && out && !out->empty()) {
            const uint32_t o0 = nd;
            uint64_t lo = 0;
            for (const pa::CopyDesc& c : *out) {
                sl.hdesc[nd++] = c;
                lo = std::max(lo, c.bytes);
            }
            hipError_t r = pa::launch_copies(sl.hdesc_dev + o0, nd - o0, lo, s);
            if (r != hipSuccess) {
                set_error(r, "launch_copies", __FILE__, __LINE__);
                rc = PA_ERR_DEVICE;
            }
        }
        hipError_t r = hipEventRecord(sl.htab_free, s);
        if (r != hipSuccess && rc == PA_OK) {
            set_error(r, "hipEventRecord", __FILE__, __LINE__);
            rc = PA_ERR_DEVICE;
        }
        return rc;
    };
    if (bflags & PA_BATCH_TUNE)
        for (auto& gr : p.groups)
            if (!gr.small && e->tuned.find(tune_key(gr)) == e->tuned.end()) {
                int rc = batch_tune_group(e, p, gr, dtab, s);
                if (rc != PA_OK) return finish(rc);
            }
    if (!(bflags & PA_BATCH_GRAPH)) return finish(batch_launch_all(e, sl, p, dtab, s));
    auto replay = [&]() -> int {
        std::string key = std::to_string((uintptr_t)sl.arena) + ":";
        for (auto& gr : p.groups)
            key += tune_key(gr) + "x" + std::to_string(gr.grid_x) + "p" + std::to_string(gr.fwd_per) + "y" +
                   std::to_string(gr.tiles_y) + ";";
        auto f = sl.graphs.find(key);
        if (f == sl.graphs.end()) {
            hipGraph_t graph = nullptr;
            PA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = batch_launch_all(e, sl, p, dtab, s);
            hipError_t ec = hipStreamEndCapture(s, &graph);
            if (rc != PA_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            PA_HIP(ec);
            hipGraphExec_t exec = nullptr;
            hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            PA_HIP(ei);
            f = sl.graphs.emplace(key, exec).first;
        }
        PA_HIP(hipGraphLaunch(f->second, s));
        return PA_OK;
    };
    return finish(replay());
}

}  // namespace

extern "C" {

int pa_batch_device(pa_engine* e, const pa_block* blocks, uint32_t count, uint32_t batch_flags, void* stream) {
    if (!e || (batch_flags & ~3u)) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    BatchPlan p;
    int rc = batch_plan(e, blocks, count, &p);
    if (rc != PA_OK) return rc;
    rc = ensure_arena(e, e->batch, p.table_bytes + p.scratch_bytes);
    if (rc != PA_OK) return rc;
    return batch_execute(e, e->batch, p, blocks, batch_flags, stream_of(e, stream));
}

int pa_run_batch(pa_engine* e, const pa_block* blocks, uint32_t count, uint32_t batch_flags) {
    if (!e || (batch_flags & ~3u)) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    BatchPlan p;
    int rc = batch_plan(e, blocks, count, &p);
    if (rc != PA_OK) return rc;
    for (uint32_t i = 0; i < count; ++i)
        if (blocks[i].margin) return PA_ERR_ARG;
    // staging after the table and scratch: per block key (n/32+2 words), seed, output
    std::vector<uint64_t> koff(count), soff(count), ooff(count);
    uint64_t off = p.table_bytes + p.scratch_bytes, key_lo = off;
    for (uint32_t i = 0; i < count; ++i) {
        koff[i] = off;
        off += align256(stage_key_bytes(p.items[i].g));
    }
    const uint64_t key_hi = off;
    for (uint32_t i = 0; i < count; ++i) {
        soff[i] = off;
        if (p.items[i].dyn) off += align256(stage_seed_bytes(p.items[i].g));
        ooff[i] = off;
        off += align256(stage_out_bytes(p.items[i].g));
    }
    rc = ensure_arena(e, e->batch, off);
    if (rc != PA_OK) return rc;
    hipStream_t s = e->stream;
    uint8_t* arena = e->batch.arena;
    std::vector<pa_block> dev(blocks, blocks + count);
    PA_HIP(hipMemsetAsync(arena + key_lo, 0, key_hi - key_lo, s));
    for (uint32_t i = 0; i < count; ++i) {
        const Geom& g = p.items[i].g;
        dev[i].key = reinterpret_cast<const uint32_t*>(arena + koff[i]);
        PA_HIP(hipMemcpyAsync(arena + koff[i], blocks[i].key, (size_t)key_words_for(g) * 4, hipMemcpyHostToDevice, s));
        if (p.items[i].dyn) {
            dev[i].seed = reinterpret_cast<const uint32_t*>(arena + soff[i]);
            PA_HIP(hipMemcpyAsync(arena + soff[i], blocks[i].seed, stage_seed_bytes(g), hipMemcpyHostToDevice, s));
        }
        dev[i].out = arena + ooff[i];
    }
    rc = batch_execute(e, e->batch, p, dev.data(), batch_flags, s);
    if (rc != PA_OK) return rc;
    for (uint32_t i = 0; i < count; ++i)
        PA_HIP(hipMemcpyAsync(blocks[i].out, arena + ooff[i], out_bytes_of(p.items[i].g), hipMemcpyDeviceToHost, s));
    PA_HIP(hipStreamSynchronize(s));
    return PA_OK;
}

static int split_run(pa_engine* e, const uint32_t* d_key, const uint32_t* d_seed, uint32_t n, uint32_t m,
                     uint32_t flags, uint32_t sub_log2, uint32_t* d_out, hipStream_t s, float* d_margin,
                     uint64_t extra_bytes, uint8_t** extra) {
    if (flags & ~3u) return PA_ERR_ARG;
    if (!(flags & PA_COMPRESS)) return PA_ERR_UNSUPPORTED;
    if (n < 64 || (n % 32) || m == 0 || (m % 32) || m >= n || n >= 0x80000000u) return PA_ERR_SIZE;
    if (sub_log2 < 6 || sub_log2 + 1 > e->cfg.max_n_log2 || sub_log2 + 1 > (uint32_t)pa::MAX_L_LOG) return PA_ERR_SIZE;
    const uint32_t b = 1u << sub_log2;
    if (b >= n) return PA_ERR_SIZE;
    const uint32_t h = n - m, wb = b / 32;
    const uint32_t nbands = (m + b - 1) / b, ncols = (h + 1 + b - 1) / b;
    // one tile's footprint: inputs, output, batch scratch and table row
    Geom tg;
    int rc = make_geom(2 * b, b, 0, e->cfg.out_reverse_endian, e->cfg.max_n_log2, &tg);
    if (rc != PA_OK) return rc;
    const uint64_t tio = align256((uint64_t)(wb + 1) * 4) + align256((uint64_t)2 * wb * 4) + align256((uint64_t)wb * 4);
    const uint64_t tile_all = tio + batch_scratch_bytes(tg, true) + sizeof(pa::BlockArgs);
    const uint64_t chunk_budget = std::max<uint64_t>(tile_all * nbands, std::min<uint64_t>(8ull << 30, e->cfg.hbm_budget_bytes / 4));
    uint32_t cols = (uint32_t)std::max<uint64_t>(1, chunk_budget / (tile_all * nbands));
    cols = std::min(cols, std::max(1u, 65535u / nbands));
    cols = std::min(cols, ncols);
    if ((uint64_t)cols * nbands > 65535u) return PA_ERR_SIZE;
    const uint32_t tiles = cols * nbands;
    // split arena: tile keys | tile seeds | tile outputs | band accumulator | caller's extra
    const uint64_t kb = align256((uint64_t)tiles * (wb + 1) * 4), sb = align256((uint64_t)tiles * 2 * wb * 4);
    const uint64_t ob = align256((uint64_t)tiles * wb * 4), ab = align256((uint64_t)nbands * wb * 4 + 16);
    const uint64_t need = kb + sb + ob + ab + align256(extra_bytes);
    if (need > e->split_cap) {
        PA_HIP(hipStreamSynchronize(s));
        if (e->split) dev_free(e, e->split, e->split_cap);
        e->split = nullptr;
        e->split_cap = 0;
        rc = dev_alloc(e, (void**)&e->split, need);
        if (rc != PA_OK) return rc;
        e->split_cap = need;
    }
    uint32_t* tkeys = reinterpret_cast<uint32_t*>(e->split);
    uint32_t* tseeds = reinterpret_cast<uint32_t*>(e->split + kb);
    uint32_t* touts = reinterpret_cast<uint32_t*>(e->split + kb + sb);
    uint32_t* acc = reinterpret_cast<uint32_t*>(e->split + kb + sb + ob);
    if (extra) {
        *extra = e->split + kb + sb + ob + ab;
        return PA_OK;  // the caller stages its buffers, then calls again with extra == nullptr
    }
    std::vector<pa_block> blocks(tiles);
    for (uint32_t t = 0; t < tiles; ++t) {
        pa_block& bl = blocks[t];
        bl = pa_block{};
        bl.key = tkeys + (size_t)t * (wb + 1);
        bl.seed = tseeds + (size_t)t * 2 * wb;
        bl.n = 2 * b;
        bl.m = b;
        bl.flags = 0;
        bl.out = reinterpret_cast<uint8_t*>(touts + (size_t)t * wb);
        bl.margin = d_margin;
    }
    for (uint32_t j0 = 0; j0 < ncols; j0 += cols) {
        const uint32_t cc = std::min(cols, ncols - j0);
        PA_HIP(pa::launch_tile_inputs(d_key, d_seed, n, h, b, nbands, j0, cc * nbands, tkeys, tseeds, s));
        BatchPlan p;
        rc = batch_plan(e, blocks.data(), cc * nbands, &p);
        if (rc != PA_OK) return rc;
        rc = ensure_arena(e, e->batch, p.table_bytes + p.scratch_bytes);
        if (rc != PA_OK) return rc;
        rc = batch_execute(e, e->batch, p, blocks.data(), PA_BATCH_GRAPH, s);
        if (rc != PA_OK) return rc;
        PA_HIP(pa::launch_tile_xor(acc, touts, nbands, wb, cc, j0 == 0, s));
    }
    PA_HIP(pa::launch_split_finish(acc, d_key, d_out, m, h, (flags & PA_XOR_KEY_REST) ? 1u : 0u,
                                   e->cfg.out_reverse_endian, s));
    return PA_OK;
}

int pa_block_device_split(pa_engine* e, const uint32_t* d_key, const uint32_t* d_seed, uint32_t n, uint32_t m,
                          uint32_t flags, uint32_t sub_log2, uint8_t* d_out, void* stream, float* d_margin) {
    if (!e || !d_key || !d_seed || !d_out) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    return split_run(e, d_key, d_seed, n, m, flags, sub_log2, reinterpret_cast<uint32_t*>(d_out),
                     stream_of(e, stream), d_margin, 0, nullptr);
}

int pa_run_block_split(pa_engine* e, const uint32_t* key, const uint32_t* seed, uint32_t n, uint32_t m,
                       uint32_t flags, uint32_t sub_log2, uint8_t* out) {
    if (!e || !key || !seed || !out) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    if (n < 64 || (n % 32) || m == 0 || (m % 32) || m >= n || n >= 0x80000000u) return PA_ERR_SIZE;
    const uint64_t kbytes = align256((uint64_t)(n / 32 + 2) * 4), sbytes = align256((uint64_t)(n / 32) * 4);
    const uint64_t obytes = align256((uint64_t)m / 8 + 16);
    hipStream_t s = e->stream;
    uint8_t* st = nullptr;
    int rc = split_run(e, nullptr, nullptr, n, m, flags, sub_log2, nullptr, s, nullptr, kbytes + sbytes + obytes, &st);
    if (rc != PA_OK) return rc;
    uint32_t* dk = reinterpret_cast<uint32_t*>(st);
    uint32_t* ds = reinterpret_cast<uint32_t*>(st + kbytes);
    uint32_t* dout = reinterpret_cast<uint32_t*>(st + kbytes + sbytes);
    const uint32_t kw = (flags & PA_XOR_KEY_REST) ? n / 32 + 1 : (n - m) / 32 + 1;
    PA_HIP(hipMemsetAsync(dk, 0, kbytes, s));
    PA_HIP(hipMemcpyAsync(dk, key, (size_t)kw * 4, hipMemcpyHostToDevice, s));
    PA_HIP(hipMemcpyAsync(ds, seed, (size_t)(n / 32) * 4, hipMemcpyHostToDevice, s));
    rc = split_run(e, dk, ds, n, m, flags, sub_log2, dout, s, nullptr, kbytes + sbytes + obytes, nullptr);
    if (rc != PA_OK) return rc;
    PA_HIP(hipMemcpyAsync(out, dout, m / 8, hipMemcpyDeviceToHost, s));
    PA_HIP(hipStreamSynchronize(s));
    return PA_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ serving ----
namespace {

using Clock = std::chrono::steady_clock;

// Serving ranges of at least this many bytes go by DMA, smaller ones through the copy kernel.
constexpr uint64_t SERVE_DMA_MIN = 1ull << 20;

// The device-mapped address of p if [p, p + bytes) lies inside one pa_host_alloc allocation,
// else null.
uint8_t* pinned_dev(pa_engine* e, const void* p, uint64_t bytes) {
    std::lock_guard<std::mutex> lk(e->pin_mu);
    auto it = e->pinned.upper_bound(reinterpret_cast<uintptr_t>(p));
    if (it == e->pinned.begin()) return nullptr;
    --it;
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    if (a < it->first || a + bytes > it->first + it->second.bytes) return nullptr;
    return it->second.dev + (a - it->first);
}

struct ServeReq {
    pa_block b{};                  // host pointers
    bool stage_out = false;        // output through the slot's host staging (caller's not pinned)
    int status = 1;                // 1 while pending, then the block's final status
    bool waited = false;           // a pa_serve_wait call holds this request
    Clock::time_point t_submit, t_done;
    uint64_t dkey = 0, dseed = 0, dout = 0;  // arena offsets
    uint64_t hout = 0;                       // host staging offset of a staged output
    size_t key_bytes = 0, seed_bytes = 0, out_bytes = 0;
};

struct ServeSlot {
    BatchSlot bs;
    bool busy = false;
    hipEvent_t done = nullptr;
    std::vector<ServeReq*> reqs;
    uint8_t* hstage = nullptr;     // pinned, mapped host staging for callers' unpinned buffers
    uint8_t* hstage_dev = nullptr;
    uint64_t hstage_cap = 0;
    uint64_t bytes = 0;            // device bytes this batch holds against the budget
};

}  // namespace

struct Server {
    pa_serve_config cfg{};
    std::thread th;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServeReq*> queue;
    std::map<pa_ticket, ServeReq*> reqs;
    std::vector<ServeSlot> slots;
    uint64_t in_use = 0;
    bool stop = false;
    int waiters = 0;  // threads inside pa_serve_wait (pa_serve_stop lets them leave first)
    pa_ticket next = 1;
    int init_status = PA_OK;
};

namespace {

// Retires finished batches: outputs from pinned staging to the callers' buffers, then completion.
bool serve_retire(Server* sv) {
    bool any = false;
    for (auto& sl : sv->slots) {
        if (!sl.busy) continue;
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        const int st = q == hipSuccess ? PA_OK : PA_ERR_DEVICE;
        if (q != hipSuccess) set_error(q, "serving batch", __FILE__, __LINE__);
        for (ServeReq* r : sl.reqs)
            if (st == PA_OK && r->stage_out) memcpy(r->b.out, sl.hstage + r->hout, r->out_bytes);
        const auto now = Clock::now();
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            for (ServeReq* r : sl.reqs) {
                r->status = st;
                r->t_done = now;
            }
            sv->in_use -= sl.bytes;
        }
        sv->cv_done.notify_all();
        sl.reqs.clear();
        sl.busy = false;
        sl.bytes = 0;
        any = true;
    }
    return any;
}

void serve_fail(Server* sv, ServeReq* r, int st) {
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        r->status = st;
        r->t_done = Clock::now();
    }
    sv->cv_done.notify_all();
}

// Device bytes a request holds while in flight: its scratch and its staging.
uint64_t serve_req_bytes(const Geom& g, bool dyn) {
    return batch_scratch_bytes(g, dyn) + align256(stage_key_bytes(g)) + (dyn ? align256(stage_seed_bytes(g)) : 0) +
           align256(stage_out_bytes(g)) + align256(sizeof(pa::BlockArgs));
}

// Forms one batch from the queue head into a free slot and launches it.  Returns false when
// nothing could be admitted.
bool serve_admit(pa_engine* e, Server* sv, ServeSlot& sl) {
    std::vector<ServeReq*> take;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        while (!sv->queue.empty() && take.size() < sv->cfg.max_batch) {
            take.push_back(sv->queue.front());
            sv->queue.pop_front();
        }
    }
    if (take.empty()) return false;
    std::lock_guard<std::mutex> elk(e->mu);
    // validate one by one; size against the budget; give back what does not fit
    // every slot's arena stays within budget / slots, so all arenas together stay within budget
    const uint64_t slot_limit = sv->cfg.hbm_budget_bytes / sv->slots.size();
    std::vector<ServeReq*> ok;
    std::vector<pa_block> blocks;
    uint64_t bytes = 0;
    const uint64_t budget_left = slot_limit - 4096;  // block table rounding
    size_t i = 0;
    for (; i < take.size(); ++i) {
        ServeReq* r = take[i];
        BatchPlan one;
        int rc = batch_plan(e, &r->b, 1, &one);
        if (rc != PA_OK) {
            serve_fail(sv, r, rc);
            continue;
        }
        const uint64_t rb = serve_req_bytes(one.items[0].g, one.items[0].dyn);
        if (rb > budget_left) {
            serve_fail(sv, r, PA_ERR_NOMEM);
            continue;
        }
        if (bytes + rb > budget_left) break;
        bytes += rb;
        ok.push_back(r);
        blocks.push_back(r->b);
    }
    if (i < take.size()) {
        std::lock_guard<std::mutex> lk(sv->mu);
        for (size_t j = take.size(); j-- > i;) sv->queue.push_front(take[j]);
    }
    if (ok.empty()) return false;
    BatchPlan p;
    int rc = batch_plan(e, blocks.data(), (uint32_t)blocks.size(), &p);
    auto fail_all = [&](int st) {
        for (ServeReq* r : ok) serve_fail(sv, r, st);
        return true;
    };
    if (rc != PA_OK) return fail_all(rc);
    // arena after table + scratch: keys, seeds, then outputs
    uint64_t off = p.table_bytes + p.scratch_bytes;
    for (size_t k = 0; k < ok.size(); ++k) {
        ServeReq* r = ok[k];
        const Geom& g = p.items[k].g;
        r->key_bytes = (size_t)key_words_for(g) * 4;  // the kernels read no key word beyond these
        r->dkey = off;
        off += align256(stage_key_bytes(g));
        r->seed_bytes = p.items[k].dyn ? stage_seed_bytes(g) : 0;
        r->dseed = off;
        off += align256(r->seed_bytes);
    }
    for (size_t k = 0; k < ok.size(); ++k) {
        ok[k]->out_bytes = out_bytes_of(p.items[k].g);
        ok[k]->dout = off;
        off += align256(stage_out_bytes(p.items[k].g));
    }
    rc = ensure_arena(e, sl.bs, off, slot_limit);
    if (rc != PA_OK) return fail_all(rc);
    // Ranges below SERVE_DMA_MIN move through the staging copy kernel, larger ones by DMA: one
    // kernel launch replaces thousands of small DMA commands, while the DMA engines move large
    // blocks alongside the other slots' kernels (the copy kernel cannot: the transform kernels
    // fill every compute unit's registers).  Either way the host side is the caller's buffer when
    // it is pinned (pa_host_alloc), else the slot's mapped host staging.
    uint8_t* A = sl.bs.arena;
    std::vector<const uint8_t*> ksrc(ok.size()), ssrc(ok.size());
    std::vector<uint8_t*> odst(ok.size());
    std::vector<uint64_t> hk(ok.size(), UINT64_MAX), hs(ok.size(), UINT64_MAX);
    uint64_t hneed = 0;
    for (size_t k = 0; k < ok.size(); ++k) {
        ServeReq* r = ok[k];
        ksrc[k] = pinned_dev(e, r->b.key, r->key_bytes);
        if (!ksrc[k]) { hk[k] = hneed; hneed += align256(r->key_bytes); }
        if (r->seed_bytes) {
            ssrc[k] = pinned_dev(e, r->b.seed, r->seed_bytes);
            if (!ssrc[k]) { hs[k] = hneed; hneed += align256(r->seed_bytes); }
        }
        odst[k] = pinned_dev(e, r->b.out, r->out_bytes);
        r->stage_out = !odst[k];
        if (r->stage_out) { r->hout = hneed; hneed += align256(r->out_bytes); }
    }
    if (hneed && host_grow(&sl.hstage, &sl.hstage_dev, &sl.hstage_cap, hneed) != hipSuccess) return fail_all(PA_ERR_NOMEM);
    // kernel ranges use device-mapped host addresses, DMA ranges the host addresses
    std::vector<pa::CopyDesc> in, outd, dma_in, dma_out;
    in.reserve(2 * ok.size());
    outd.reserve(ok.size());
    std::vector<pa_block> dev(blocks);
    auto add_in = [&](const void* host, const uint8_t* mapped, uint8_t* d, uint64_t bytes) {
        if (bytes >= SERVE_DMA_MIN) dma_in.push_back({static_cast<const uint8_t*>(host), d, bytes});
        else in.push_back({mapped, d, bytes});
    };
    for (size_t k = 0; k < ok.size(); ++k) {
        ServeReq* r = ok[k];
        const void* kh = r->b.key;
        if (!ksrc[k]) {
            memcpy(sl.hstage + hk[k], r->b.key, r->key_bytes);
            kh = sl.hstage + hk[k];
            ksrc[k] = sl.hstage_dev + hk[k];
        }
        add_in(kh, ksrc[k], A + r->dkey, r->key_bytes);
        dev[k].key = reinterpret_cast<const uint32_t*>(A + r->dkey);
        if (r->seed_bytes) {
            const void* sh = r->b.seed;
            if (!ssrc[k]) {
                memcpy(sl.hstage + hs[k], r->b.seed, r->seed_bytes);
                sh = sl.hstage + hs[k];
                ssrc[k] = sl.hstage_dev + hs[k];
            }
            add_in(sh, ssrc[k], A + r->dseed, r->seed_bytes);
            dev[k].seed = reinterpret_cast<const uint32_t*>(A + r->dseed);
        }
        if (r->out_bytes >= SERVE_DMA_MIN)
            dma_out.push_back({A + r->dout, r->stage_out ? sl.hstage + r->hout : r->b.out, r->out_bytes});
        else
            outd.push_back({A + r->dout, r->stage_out ? sl.hstage_dev + r->hout : odst[k], r->out_bytes});
        dev[k].out = A + r->dout;
        dev[k].margin = nullptr;
    }
    hipStream_t s = sl.bs.stream;
    auto launch = [&]() -> int {
        for (const pa::CopyDesc& c : dma_in) PA_HIP(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyHostToDevice, s));
        int r2 = batch_execute(e, sl.bs, p, dev.data(), sv->cfg.batch_flags, s, &in, &outd);
        if (r2 != PA_OK) return r2;
        for (const pa::CopyDesc& c : dma_out) PA_HIP(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, s));
        PA_HIP(hipEventRecord(sl.done, s));
        return PA_OK;
    };
    rc = launch();
    if (rc != PA_OK) return fail_all(rc);
    sl.reqs = ok;
    sl.bytes = bytes;
    sl.busy = true;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->in_use += bytes;
    }
    return true;
}

void serve_loop(pa_engine* e, Server* sv) {
    (void)hipSetDevice(e->cfg.device);
    while (true) {
        bool progressed = serve_retire(sv);
        for (auto& sl : sv->slots)
            if (!sl.busy && serve_admit(e, sv, sl)) progressed = true;
        bool busy = false;
        for (auto& sl : sv->slots) busy = busy || sl.busy;
        std::unique_lock<std::mutex> lk(sv->mu);
        if (sv->stop && sv->queue.empty() && !busy) break;
        if (progressed) continue;
        if (busy) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        } else {
            sv->cv_work.wait_for(lk, std::chrono::milliseconds(2),
                                 [&] { return sv->stop || !sv->queue.empty(); });
        }
    }
}

}  // namespace

extern "C" {

int pa_host_alloc(pa_engine* e, uint64_t bytes, void** ptr) {
    if (!e || !ptr || bytes == 0) return PA_ERR_ARG;
    void* p = nullptr;
    uint8_t* d = nullptr;
    hipError_t r = host_alloc_mapped(&p, &d, bytes);
    if (r != hipSuccess) {
        set_error(r, "hipHostMalloc", __FILE__, __LINE__);
        return PA_ERR_NOMEM;
    }
    std::lock_guard<std::mutex> lk(e->pin_mu);
    e->pinned[reinterpret_cast<uintptr_t>(p)] = {bytes, d};
    *ptr = p;
    return PA_OK;
}

int pa_host_free(pa_engine* e, void* p) {
    if (!e || !p) return PA_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(e->pin_mu);
        auto it = e->pinned.find(reinterpret_cast<uintptr_t>(p));
        if (it == e->pinned.end()) return PA_ERR_NOT_FOUND;
        e->pinned.erase(it);
    }
    PA_HIP(hipHostFree(p));
    return PA_OK;
}

int pa_serve_start(pa_engine* e, const pa_serve_config* cfg) {
    if (!e) return PA_ERR_ARG;
    if (e->server) return PA_ERR_ARG;
    auto sv = std::make_unique<Server>();
    sv->cfg.slots = 3;
    sv->cfg.max_batch = 256;
    sv->cfg.hbm_budget_bytes = 16ull << 30;
    if (cfg) {
        if (cfg->reserved || (cfg->batch_flags & ~3u)) return PA_ERR_ARG;
        if (cfg->slots) sv->cfg.slots = cfg->slots;
        if (cfg->max_batch) sv->cfg.max_batch = std::min<uint32_t>(cfg->max_batch, 65535u);
        if (cfg->hbm_budget_bytes) sv->cfg.hbm_budget_bytes = cfg->hbm_budget_bytes;
        sv->cfg.batch_flags = cfg->batch_flags;
    }
    if (sv->cfg.slots > 16) return PA_ERR_ARG;
    sv->slots.resize(sv->cfg.slots);
    for (auto& sl : sv->slots) {
        int rc = slot_init(&sl.bs, true);
        if (rc != PA_OK) return rc;
        PA_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    }
    {
        // Each slot's arena starts at a quarter of its share of the budget (at most 4 GiB, plus the
        // usual half again), so that the first batches do not stop the stream to allocate; it
        // still grows on demand up to the share.  Best effort: a failure leaves on-demand growth.
        std::lock_guard<std::mutex> lk(e->mu);
        const uint64_t share = sv->cfg.hbm_budget_bytes / sv->slots.size();
        const uint64_t want = std::min<uint64_t>(share / 4, 4ull << 30);
        for (auto& sl : sv->slots)
            if (ensure_arena(e, sl.bs, want, share) != PA_OK) break;
    }
    Server* raw = sv.release();
    e->server = raw;
    raw->th = std::thread(serve_loop, e, raw);
    return PA_OK;
}

int pa_serve_submit_many(pa_engine* e, const pa_block* blocks, uint32_t count, pa_ticket* tickets) {
    if (!e || !e->server || !blocks || !tickets) return PA_ERR_ARG;
    for (uint32_t i = 0; i < count; ++i) {
        const pa_block* b = blocks + i;
        if (!b->key || !b->out || (!b->seed && !b->seed_id) || b->margin) return PA_ERR_ARG;
    }
    Server* sv = e->server;
    std::vector<ServeReq*> rs(count);
    const auto now = Clock::now();
    for (uint32_t i = 0; i < count; ++i) {
        rs[i] = new ServeReq();
        rs[i]->b = blocks[i];
        rs[i]->t_submit = now;
    }
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->stop) {
            for (ServeReq* r : rs) delete r;
            return PA_ERR_ARG;
        }
        for (uint32_t i = 0; i < count; ++i) {
            tickets[i] = sv->next++;
            sv->reqs[tickets[i]] = rs[i];
            sv->queue.push_back(rs[i]);
        }
    }
    sv->cv_work.notify_one();
    return PA_OK;
}

int pa_serve_submit(pa_engine* e, const pa_block* b, pa_ticket* ticket) {
    if (!b || !ticket) return PA_ERR_ARG;
    return pa_serve_submit_many(e, b, 1, ticket);
}

int pa_serve_wait(pa_engine* e, pa_ticket ticket, int64_t timeout_ns, uint64_t* latency_ns) {
    if (!e || !e->server) return PA_ERR_ARG;
    Server* sv = e->server;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return PA_ERR_NOT_FOUND;
    ServeReq* r = it->second;
    if (r->waited) return PA_ERR_ARG;  // one waiter per ticket: the winner frees the request
    r->waited = true;
    ++sv->waiters;
    auto leave = [&](int st) {
        --sv->waiters;
        sv->cv_done.notify_all();
        return st;
    };
    auto ready = [&] { return r->status != 1; };
    if (timeout_ns < 0) sv->cv_done.wait(lk, ready);
    else if (!sv->cv_done.wait_for(lk, std::chrono::nanoseconds(timeout_ns), ready)) {
        r->waited = false;
        return leave(PA_ERR_TIMEOUT);
    }
    const int st = r->status;
    if (latency_ns)
        *latency_ns = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(r->t_done - r->t_submit).count();
    sv->reqs.erase(it);
    delete r;
    return leave(st);
}

int pa_serve_stop(pa_engine* e) {
    if (!e || !e->server) return PA_ERR_ARG;
    Server* sv = e->server;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->th.joinable()) sv->th.join();
    {
        // every request has its final status now: let threads already inside pa_serve_wait leave
        std::unique_lock<std::mutex> lk(sv->mu);
        sv->cv_done.notify_all();
        sv->cv_done.wait(lk, [&] { return sv->waiters == 0; });
    }
    for (auto& sl : sv->slots) {
        slot_release(e, &sl.bs);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.hstage) (void)hipHostFree(sl.hstage);
    }
    for (auto& kv : sv->reqs) delete kv.second;
    delete sv;
    e->server = nullptr;
    return PA_OK;
}

int pa_stage_timing(pa_engine* e, int enable) {
    if (!e) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    e->timing = enable != 0;
    return PA_OK;
}

int pa_stage_times(pa_engine* e, const char** names, double* ms, uint32_t* counts, uint32_t* n_stages) {
    if (!e || !n_stages) return PA_ERR_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    for (auto& p : e->pending) {
        PA_HIP(hipEventSynchronize(p.b));
        float dt = 0.f;
        PA_HIP(hipEventElapsedTime(&dt, p.a, p.b));
        e->stage_ms[p.stage] += dt;
        e->stage_count[p.stage] += 1;
        e->event_pool.push_back(p.a);
        e->event_pool.push_back(p.b);
    }
    e->pending.clear();
    uint32_t cap = *n_stages, k = 0;
    for (int st = 0; st < ST_COUNT; ++st) {
        if (e->stage_count[st] == 0) continue;
        if (k < cap) {
            if (names) names[k] = kStageNames[st];
            if (ms) ms[k] = e->stage_ms[st];
            if (counts) counts[k] = e->stage_count[st];
        }
        ++k;
        e->stage_ms[st] = 0.0;
        e->stage_count[st] = 0;
    }
    *n_stages = k < cap ? k : cap;
    return PA_OK;
}

int pa_synth_bits(uint64_t seed, double p, uint64_t nbits, uint32_t* out) {
    if (!out || p < 0.0 || p > 1.0) return PA_ERR_ARG;
    auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
    uint64_t sm = seed, s[4];
    for (int i = 0; i < 4; ++i) {
        uint64_t z = (sm += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        s[i] = z ^ (z >> 31);
    }
    auto next = [&]() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    };
    const uint64_t nw = (nbits + 31) / 32;
    if (p == 0.5) {
        for (uint64_t i = 0; i < nw; i += 2) {
            const uint64_t r = next();
            out[i] = (uint32_t)(r >> 32);
            if (i + 1 < nw) out[i + 1] = (uint32_t)r;
        }
    } else {
        memset(out, 0, nw * 4);
        for (uint64_t b = 0; b < nbits; ++b)
            if ((double)(next() >> 11) * 0x1.0p-53 < p) out[b >> 5] |= 1u << (31 - (b & 31));
    }
    if (nbits & 31) out[nw - 1] &= ~((1u << (32 - (nbits & 31))) - 1u);
    return PA_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ self-tests ----
// The counterparts of the reference's unitTest* argv modes (PrivacyAmplification.cu:299-737,
// dispatched at :1645-1652).  The reference tests each of its five small kernels on closed-form
// inputs; this engine fuses those steps into its passes, so each test runs the production device
// code of that step -- the same device functions or the production kernels themselves -- on the
// same closed-form inputs and checks the result on the host.
namespace {

struct DevMem {
    void* p = nullptr;
    ~DevMem() { if (p) (void)hipFree(p); }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};
#define ST_ALLOC(buf, bytes) PA_HIP(hipMalloc(&(buf).p, (bytes)))

struct Report {
    char* out;
    uint32_t cap;
    bool ok = true;
    void fail(const char* fmt, ...) __attribute__((format(printf, 2, 3))) {
        if (ok && out && cap) {
            va_list ap;
            va_start(ap, fmt);
            vsnprintf(out, cap, fmt, ap);
            va_end(ap);
        }
        ok = false;
    }
    int done(const char* what) {
        if (ok && out && cap) snprintf(out, cap, "%s: passed", what);
        return ok ? PA_OK : PA_ERR_CHECK;
    }
};

// Ones among the binary digits of 0, 1, ..., x (OEIS A000788, the popcount closed form of
// unitTestBinInt2float): per bit b, full periods of 2^(b+1) contribute 2^b each, and the last
// partial period max(0, rem - 2^b).
uint64_t ones_upto(uint64_t x) {
    uint64_t v = 0;
    for (int b = 0; b < 40; ++b) {
        const uint64_t period = 2ull << b, half = 1ull << b;
        const uint64_t full = (x + 1) / period, rem = (x + 1) % period;
        v += full * half + (rem > half ? rem - half : 0);
    }
    return v;
}

int st_correction(pa_engine* e, Report& rp) {
    // i, j < 64 with L = 64, then 4096 pairs in [2^25, 2^27] with L = 2^27 (:299-361)
    std::vector<uint32_t> wk, ws, Ls;
    for (uint32_t i = 0; i < 64; ++i)
        for (uint32_t j = 0; j < 64; ++j) { ws.push_back(i); wk.push_back(j); Ls.push_back(64); }
    uint64_t x = 777;
    for (int k = 0; k < 2 * 4096; ++k) {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        const uint32_t v = (1u << 25) + (uint32_t)((x >> 33) % ((1u << 27) - (1u << 25) + 1u));
        if (k & 1) { wk.push_back(v); Ls.push_back(1u << 27); } else ws.push_back(v);
    }
    const uint32_t cnt = (uint32_t)wk.size();
    DevMem dwk, dws, dout;
    ST_ALLOC(dwk, cnt * 4);
    ST_ALLOC(dws, cnt * 4);
    ST_ALLOC(dout, cnt * 4);
    PA_HIP(hipMemcpy(dwk.p, wk.data(), cnt * 4, hipMemcpyHostToDevice));
    PA_HIP(hipMemcpy(dws.p, ws.data(), cnt * 4, hipMemcpyHostToDevice));
    std::vector<float> got(cnt);
    for (uint32_t lo = 0; lo < cnt;) {  // one launch per L
        uint32_t hi = lo;
        while (hi < cnt && Ls[hi] == Ls[lo]) ++hi;
        PA_HIP(pa::launch_st_correction(dwk.as<uint32_t>() + lo, dws.as<uint32_t>() + lo, Ls[lo], dout.as<float>() + lo,
                                        hi - lo, e->stream));
        lo = hi;
    }
    PA_HIP(hipMemcpyAsync(got.data(), dout.p, cnt * 4, hipMemcpyDeviceToHost, e->stream));
    PA_HIP(hipStreamSynchronize(e->stream));
    for (uint32_t k = 0; k < cnt; ++k) {
        // the reference's host expectation: the 32-bit product, / n, mod 2
        const double want = std::fmod((double)(uint32_t)(wk[k] * ws[k]) / (double)Ls[k], 2.0);
        if (std::fabs(got[k] - want) > 1e-4) {
            rp.fail("CalculateCorrectionFloat: case %u (wk %u, ws %u, L %u): %.7f, want %.7f", k, wk[k], ws[k], Ls[k],
                    got[k], want);
            break;
        }
    }
    return rp.done("CalculateCorrectionFloat");
}

// The seed's split spectrum through the production pass A and row pass (pa_seed_put's launches).
int st_seed_spectrum(pa_engine* e, const uint32_t* h_seed, uint32_t n, std::vector<float2>& S, uint32_t& ws) {
    Geom g;
    Workspace* w;
    int rc = prepare(e, n, 3 * n / 8, PA_DEFAULT_FLAGS, &g, &w);
    if (rc != PA_OK) return rc;
    DevMem dS, dws, dseed;
    ST_ALLOC(dS, (size_t)(g.N + 1) * 8);
    ST_ALLOC(dws, pa::POP_SLOTS * 4);
    ST_ALLOC(dseed, (size_t)n / 8);
    hipStream_t s = e->stream;
    PA_HIP(hipMemcpyAsync(dseed.p, h_seed, n / 8, hipMemcpyHostToDevice, s));
    const pa::Tables t = tables(e, g.L);
    PA_HIP(pa::launch_zero_u32(dws.as<uint32_t>(), pa::POP_SLOTS, s));
    PA_HIP(pa::launch_transpose(dseed.as<uint32_t>(), w->tbs, dws.as<uint32_t>(), g, 1, s));
    PA_HIP(pa::launch_fwd_cols(w->tbs, dS.as<float2>(), g, t, s));
    PA_HIP(pa::launch_rows(pa::ROWS_SEEDPREP, nullptr, dS.as<float2>(), nullptr, g, t, e->cus, s));
    S.resize(g.N + 1);
    std::vector<uint32_t> part(pa::POP_SLOTS);
    PA_HIP(hipMemcpyAsync(S.data(), dS.p, (size_t)(g.N + 1) * 8, hipMemcpyDeviceToHost, s));
    PA_HIP(hipMemcpyAsync(part.data(), dws.p, pa::POP_SLOTS * 4, hipMemcpyDeviceToHost, s));
    PA_HIP(hipStreamSynchronize(s));
    ws = 0;
    for (uint32_t v : part) ws += v;
    return PA_OK;
}

int st_dc_zero(pa_engine* e, Report& rp) {
    // setFirstElementToZero (:377-418): the DC bin of the seed spectrum leaves the row pass as an
    // exact zero; an all-ones seed (a constant sequence, spectrum at DC only) leaves nothing else
    const uint32_t n = 1u << 20;
    std::vector<uint32_t> ones(n / 32, 0xFFFFFFFFu), rnd(n / 32);
    pa_synth_bits(0x5E1F, 0.5, n, rnd.data());
    for (int which = 0; which < 2 && rp.ok; ++which) {
        std::vector<float2> S;
        uint32_t ws = 0;
        int rc = st_seed_spectrum(e, which ? rnd.data() : ones.data(), n, S, ws);
        if (rc != PA_OK) return rc;
        uint32_t want_ws = 0;
        for (uint32_t v : (which ? rnd : ones)) want_ws += (uint32_t)__builtin_popcount(v);
        if (S[0].x != 0.f || S[0].y != 0.f) rp.fail("SetFirstElementToZero: DC bin (%g, %g), want 0", S[0].x, S[0].y);
        else if (ws != want_ws) rp.fail("SetFirstElementToZero: seed popcount %u, want %u", ws, want_ws);
        double mx = 0.0;
        for (size_t k = 1; k < S.size(); ++k) mx = std::max(mx, (double)std::hypot(S[k].x, S[k].y));
        if (rp.ok && which == 0 && mx > 1e-4 * n)
            rp.fail("SetFirstElementToZero: constant seed leaves |S| %.3g outside DC", mx);
        if (rp.ok && which == 1 && mx < 1.0) rp.fail("SetFirstElementToZero: random seed spectrum vanished");
    }
    return rp.done("SetFirstElementToZero");
}

int st_product(pa_engine* e, Report& rp) {
    // ElementWiseProduct (:422-481): do1 = i + 0.77, do2 = i + 0.88 over 2^10 complex values, r = 2^5
    const uint32_t N = 1024;
    const float r = 32.f;
    std::vector<float2> a(N), b(N), z(N), prod(N), zz(N);
    for (uint32_t i = 0; i < N; ++i) {
        a[i] = make_float2(2 * i + 0.77f, 2 * i + 1 + 0.77f);
        b[i] = make_float2(2 * i + 0.88f, 2 * i + 1 + 0.88f);
        z[i] = make_float2(std::sin(0.37 * i + 0.1), std::cos(1.13 * i + 0.2));
    }
    DevMem da, db, dz, dp, dzz;
    for (DevMem* d : {&da, &db, &dz, &dp, &dzz}) ST_ALLOC(*d, N * 8);
    PA_HIP(hipMemcpy(da.p, a.data(), N * 8, hipMemcpyHostToDevice));
    PA_HIP(hipMemcpy(db.p, b.data(), N * 8, hipMemcpyHostToDevice));
    PA_HIP(hipMemcpy(dz.p, z.data(), N * 8, hipMemcpyHostToDevice));
    PA_HIP(pa::launch_st_product(da.as<float2>(), db.as<float2>(), r, dp.as<float2>(), dz.as<float2>(),
                                 dzz.as<float2>(), N, e->stream));
    PA_HIP(hipMemcpyAsync(prod.data(), dp.p, N * 8, hipMemcpyDeviceToHost, e->stream));
    PA_HIP(hipMemcpyAsync(zz.data(), dzz.p, N * 8, hipMemcpyDeviceToHost, e->stream));
    PA_HIP(hipStreamSynchronize(e->stream));
    for (uint32_t i = 0; i < N && rp.ok; ++i) {
        const double ar = (2.0 * i + 0.77) / r, ai = (2.0 * i + 1 + 0.77) / r;
        const double br = (2.0 * i + 0.88) / r, bi = (2.0 * i + 1 + 0.88) / r;
        const double re = ar * br - ai * bi, im = ar * bi + ai * br;
        if (std::fabs(prod[i].x - re) > 1e-3 || std::fabs(prod[i].y - im) > 1e-3)
            rp.fail("ElementWiseProduct: value %u (%.6f, %.6f), want (%.6f, %.6f)", i, prod[i].x, prod[i].y, re, im);
    }
    for (uint32_t i = 1; i < N && rp.ok; ++i) {
        if (i == N / 2) continue;
        if (std::hypot(zz[i].x - 4 * z[i].x, zz[i].y - 4 * z[i].y) > 1e-5)
            rp.fail("ElementWiseProduct: split round trip at %u (%.7f, %.7f), want 4 z = (%.7f, %.7f)", i, zz[i].x,
                    zz[i].y, 4 * z[i].x, 4 * z[i].y);
    }
    return rp.done("ElementWiseProduct");
}

inline uint32_t refbit(const uint32_t* w, uint64_t p) { return (w[p >> 5] >> (31u - (p & 31u))) & 1u; }

int st_unpack(pa_engine* e, Report& rp) {
    // binInt2float (:484-602): binIn[i] = i; the popcount of n bits is A000788(n/32 - 1); here the
    // production bit transpose of pass A also lays the bits out for the class tables, checked too
    const uint32_t nmax = 1u << 27;
    std::vector<uint32_t> bin(nmax / 32);
    for (uint32_t i = 0; i < nmax / 32; ++i) bin[i] = i;
    DevMem dbin, dcnt;
    ST_ALLOC(dbin, nmax / 8);
    ST_ALLOC(dcnt, pa::POP_SLOTS * 4);
    PA_HIP(hipMemcpy(dbin.p, bin.data(), nmax / 8, hipMemcpyHostToDevice));
    uint64_t lcg = 12345;
    for (uint32_t lg = pa::SMALL_MAX_LOG_N + 2; lg <= 27 && rp.ok; ++lg) {
        const uint32_t n = 1u << lg;
        Geom g;
        Workspace* w;
        int rc = prepare(e, n, 3 * n / 8, PA_DEFAULT_FLAGS, &g, &w);
        if (rc != PA_OK) return rc;
        PA_HIP(pa::launch_zero_u32(dcnt.as<uint32_t>(), pa::POP_SLOTS, e->stream));
        PA_HIP(pa::launch_transpose(dbin.as<uint32_t>(), w->tbs, dcnt.as<uint32_t>(), g, 1, e->stream));
        std::vector<uint32_t> part(pa::POP_SLOTS), tb(g.N / g.D);
        PA_HIP(hipMemcpyAsync(part.data(), dcnt.p, pa::POP_SLOTS * 4, hipMemcpyDeviceToHost, e->stream));
        PA_HIP(hipMemcpyAsync(tb.data(), w->tbs, tb.size() * 4, hipMemcpyDeviceToHost, e->stream));
        PA_HIP(hipStreamSynchronize(e->stream));
        uint64_t cnt = 0;
        for (uint32_t v : part) cnt += v;
        if (cnt != ones_upto(n / 32 - 1)) {
            rp.fail("BinInt2float: n = 2^%u popcount %llu, want %llu", lg, (unsigned long long)cnt,
                    (unsigned long long)ones_upto(n / 32 - 1));
            break;
        }
        // tb[u N2 + t2] = E | O << 16: bit v of E / O = seed bit 2(t2 + N2 (u + M v)) / that + 1
        const uint64_t words = tb.size(), checks = lg <= 21 ? words : 65536;
        for (uint64_t c = 0; c < checks && rp.ok; ++c) {
            uint64_t idx = c;
            if (checks != words) {
                lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
                idx = (lcg >> 20) % words;
            }
            const uint64_t u = idx / g.N2, t2 = idx % g.N2;
            uint32_t want = 0;
            for (uint32_t v = 0; v < g.D; ++v) {
                const uint64_t p = 2 * (t2 + (uint64_t)g.N2 * (u + (uint64_t)g.M * v));
                want |= refbit(bin.data(), p) << v;
                want |= refbit(bin.data(), p + 1) << (16 + v);
            }
            if (tb[idx] != want) rp.fail("BinInt2float: n = 2^%u bit word %llu = %08x, want %08x", lg,
                                         (unsigned long long)idx, tb[idx], want);
        }
    }
    return rp.done("BinInt2float");
}

int st_epilogue(pa_engine* e, Report& rp) {
    // ToBinaryArray (:605-737): v[p] = bit 31 - p % 32 of p / 32, corr = 1.9, norm = 1, key rest
    // 0xAAAAAAAA words; output bit p = v[p] ^ key-rest bit p (NoXOR: v[p]), n = 2^10 .. 2^27,
    // m = n/4 + n/8.  Pass A''s rounding (round_bit) writes the nibbles, the production pack stage
    // assembles, XORs and orders the words.
    const uint32_t nmax = 1u << 27;
    DevMem dkey, dnib, dout;
    ST_ALLOC(dkey, (nmax / 32 + 2) * 4);
    ST_ALLOC(dout, nmax / 8 + 64);
    ST_ALLOC(dnib, (size_t)nmax / 2 + 64);
    {
        // key words 0x55555555: key rest word w = (key[h/32+w] << 1) | (key[h/32+w+1] >> 31) = 0xAAAAAAAA
        std::vector<uint32_t> key(nmax / 32 + 2, 0x55555555u);
        PA_HIP(hipMemcpy(dkey.p, key.data(), key.size() * 4, hipMemcpyHostToDevice));
    }
    for (uint32_t lg = pa::SMALL_MAX_LOG_N + 2; lg <= 27 && rp.ok; ++lg) {
        const uint32_t n = 1u << lg, m = n / 4 + n / 8;
        for (int do_xor = 1; do_xor >= 0 && rp.ok; --do_xor) {
            Geom g;
            int rc = make_geom(n, m, PA_COMPRESS | (do_xor ? PA_XOR_KEY_REST : 0u), e->cfg.out_reverse_endian,
                               e->cfg.max_n_log2, &g);
            if (rc != PA_OK) return rc;
            PA_HIP(pa::launch_st_nibbles(dnib.as<uint8_t>(), g, 1.9f, 1.0f, e->stream));
            PA_HIP(pa::launch_pack(dnib.as<uint8_t>(), dkey.as<uint32_t>(), dout.as<uint32_t>(), g, e->stream));
            std::vector<uint8_t> out(m / 8);
            PA_HIP(hipMemcpyAsync(out.data(), dout.p, m / 8, hipMemcpyDeviceToHost, e->stream));
            PA_HIP(hipStreamSynchronize(e->stream));
            for (uint32_t p = 0; p < m; ++p) {
                const uint32_t v = ((p >> 5) >> (31u - (p & 31u))) & 1u;
                const uint32_t want = v ^ (do_xor ? (0xAAAAAAAAu >> (31u - (p & 31u))) & 1u : 0u);
                uint32_t got;
                if (g.reverse_endian) got = (out[p >> 3] >> (7u - (p & 7u))) & 1u;
                else got = refbit(reinterpret_cast<const uint32_t*>(out.data()), p);
                if (got != want) {
                    rp.fail("ToBinaryArray%s: n = 2^%u output bit %u = %u, want %u", do_xor ? "" : "NoXOR", lg, p, got, want);
                    break;
                }
            }
        }
    }
    return rp.done("ToBinaryArray");
}

}  // namespace

extern "C" {

int pa_selftest(pa_engine* e, const char* name, char* report, uint32_t report_bytes) {
    if (!e || !name) return PA_ERR_ARG;
    if (report && report_bytes) report[0] = '\0';
    std::lock_guard<std::mutex> lk(e->mu);
    PA_HIP(hipSetDevice(e->cfg.device));
    Report rp{report, report_bytes};
    const std::string t = name;
    if (t == "CalculateCorrectionFloat") return st_correction(e, rp);
    if (t == "SetFirstElementToZero") return st_dc_zero(e, rp);
    if (t == "ElementWiseProduct") return st_product(e, rp);
    if (t == "BinInt2float") return st_unpack(e, rp);
    if (t == "ToBinaryArray") return st_epilogue(e, rp);
    return PA_ERR_ARG;
}

const char* pa_strerror(int status) {
    switch (status) {
    case PA_OK: return "ok";
    case PA_ERR_ARG: return "invalid argument";
    case PA_ERR_SIZE: return g_size_error[0] ? g_size_error : "block size outside the supported family";
    case PA_ERR_DEVICE: return g_last_error[0] ? g_last_error : "HIP runtime error";
    case PA_ERR_NOMEM: return "device memory or HBM budget exhausted";
    case PA_ERR_TIMEOUT: return "timed out";
    case PA_ERR_NOT_FOUND: return "unknown seed or ticket";
    case PA_ERR_UNSUPPORTED: return "unsupported request";
    case PA_ERR_CHECK: return "self-test failed";
    default: return "unknown status";
    }
}

const char* pa_version(void) { return "pa-mi355x 0.1 (gfx950)"; }

}  // extern "C"

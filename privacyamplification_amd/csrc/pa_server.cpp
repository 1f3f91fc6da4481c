// pa_server.cpp -- the ZeroMQ front-end: a drop-in process for the reference's PrivacyAmplification
// binary around the MI355X engine.
//
// It reads the reference's config.yaml fields (readConfig, PrivacyAmplification.cu:1228-1272) and
// keeps its three sockets and frame layouts:
//   seed-in  PULL, connects to address_seed_in: [int32 reuse_seed_amount][n/8 bytes seed]
//            (reciveDataSeed, :980-1014)
//   key-in   PULL, connects to address_key_in: [u8 do_xor_key_rest][u8 do_compress]
//            [u32 vertical_block][key words] (reciveDataKey, :1036-1073)
//   amp-out  PUSH, binds address_amp_out: [m/8 bytes, or n/16 without compression] (sendData,
//            :1143-1200)
// with its seed reuse rule (a seed with reuse k > 0 serves k blocks, 0 one block, -1 every later one; mainloop
// :1779-1791, 2099-2101), its file inputs when a server is off (keyfile_path, toeplitz_seed_path),
// ampout.bin for the first store_first_ampouts_in_file outputs, verify_ampout (SHA3-256 against an
// expected digest, on a pool of verify_ampout_threads workers) and the Blocktime line.  Exit codes as the reference: 101 verification failed,
// 102 config, 103/104 seed file missing/short, 105/106 key file missing/short, 400 bad factor_exp.
//
// Blocks go through the engine's serving scheduler (pa_serve_*) from pinned buffers
// (pa_host_alloc); resident seeds (pa_seed_put) serve reused seeds; outputs leave in block order.
//
//   pa_server [--config config.yaml] [--factor_exp N] [--blocks K] [speedtest | unitTest<Name>]
// (--factor_exp overrides the config; --blocks K ends the process after K blocks.)
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pa_engine.h"
#include "sha3.hpp"
#include "zmtp.hpp"

namespace {

struct Config {
    uint32_t factor_exp = 27, reduction_exp = 11, pre_mul_reduction_exp = 5, gpu_device_id_to_use = 0;
    uint32_t input_blocks_to_cache = 16, output_blocks_to_cache = 16, verify_ampout_threads = 8;
    int32_t reuse_seed_amount = 0, vertical_len = 50331648, show_ampout = 0, store_first_ampouts_in_file = 1;
    bool do_xor_key_rest = true, do_compress = true, show_zeromq_status = true;
    bool use_matrix_seed_server = true, use_key_server = true, host_ampout_server = true, verify_ampout = true;
    std::string address_seed_in = "tcp://127.0.0.1:45555", address_key_in = "tcp://127.0.0.1:47777";
    std::string address_amp_out = "tcp://127.0.0.1:48888";
    std::string toeplitz_seed_path = "toeplitz_seed.bin", keyfile_path = "keyfile.bin";
    // SHA3-256 that verify_ampout expects; default: the reference's digest of its 2^27 test block
    // (ampout_sha3, PrivacyAmplification.h:170-173).
    std::string verify_ampout_sha3 = "c422b6865c72cad82cc26a1462b8a4566f911750f31b1475691269c1b7d4a716";
};

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// The flat "key: value" subset of YAML the reference's config.yaml uses ('#' comments, quoted
// strings, true/false).  Unknown keys are ignored, as the reference ignores them.
bool load_config(const std::string& path, Config& c, std::string& err) {
    std::ifstream f(path);
    if (!f) {
        err = "can't open " + path;
        return false;
    }
    std::string line;
    int ln = 0;
    while (std::getline(f, line)) {
        ++ln;
        std::string t = line;
        bool q1 = false, q2 = false;
        for (size_t i = 0; i < t.size(); ++i) {  // strip a comment outside quotes
            if (t[i] == '\'' && !q2) q1 = !q1;
            else if (t[i] == '"' && !q1) q2 = !q2;
            else if (t[i] == '#' && !q1 && !q2) {
                t.resize(i);
                break;
            }
        }
        t = trim(t);
        if (t.empty()) continue;
        const size_t colon = t.find(':');
        if (colon == std::string::npos) {
            err = path + ":" + std::to_string(ln) + ": expected key: value";
            return false;
        }
        const std::string k = trim(t.substr(0, colon));
        std::string v = trim(t.substr(colon + 1));
        if (v.size() >= 2 && ((v.front() == '\'' && v.back() == '\'') || (v.front() == '"' && v.back() == '"')))
            v = v.substr(1, v.size() - 2);
        auto as_bool = [&](bool& out) {
            if (v == "true" || v == "True" || v == "1") out = true;
            else if (v == "false" || v == "False" || v == "0") out = false;
            else return false;
            return true;
        };
        auto as_int = [&](auto& out) {
            char* end = nullptr;
            const long long x = std::strtoll(v.c_str(), &end, 10);
            if (v.empty() || *end) return false;
            out = static_cast<std::remove_reference_t<decltype(out)>>(x);
            return true;
        };
        bool ok = true;
        if (k == "factor_exp") ok = as_int(c.factor_exp);
        else if (k == "reuse_seed_amount") ok = as_int(c.reuse_seed_amount);
        else if (k == "vertical_len") ok = as_int(c.vertical_len);
        else if (k == "do_xor_key_rest") ok = as_bool(c.do_xor_key_rest);
        else if (k == "do_compress") ok = as_bool(c.do_compress);
        else if (k == "reduction_exp") ok = as_int(c.reduction_exp);
        else if (k == "pre_mul_reduction_exp") ok = as_int(c.pre_mul_reduction_exp);
        else if (k == "gpu_device_id_to_use") ok = as_int(c.gpu_device_id_to_use);
        else if (k == "input_blocks_to_cache") ok = as_int(c.input_blocks_to_cache);
        else if (k == "output_blocks_to_cache") ok = as_int(c.output_blocks_to_cache);
        else if (k == "show_ampout") ok = as_int(c.show_ampout);
        else if (k == "show_zeromq_status") ok = as_bool(c.show_zeromq_status);
        else if (k == "use_matrix_seed_server") ok = as_bool(c.use_matrix_seed_server);
        else if (k == "address_seed_in") c.address_seed_in = v;
        else if (k == "toeplitz_seed_path") c.toeplitz_seed_path = v;
        else if (k == "use_key_server") ok = as_bool(c.use_key_server);
        else if (k == "address_key_in") c.address_key_in = v;
        else if (k == "keyfile_path") c.keyfile_path = v;
        else if (k == "host_ampout_server") ok = as_bool(c.host_ampout_server);
        else if (k == "address_amp_out") c.address_amp_out = v;
        else if (k == "store_first_ampouts_in_file") {
            bool b;
            ok = as_int(c.store_first_ampouts_in_file) || (as_bool(b) && ((c.store_first_ampouts_in_file = b), true));
        } else if (k == "verify_ampout") ok = as_bool(c.verify_ampout);
        else if (k == "verify_ampout_threads") ok = as_int(c.verify_ampout_threads);
        else if (k == "verify_ampout_sha3") c.verify_ampout_sha3 = v;
        if (!ok) {
            err = path + ":" + std::to_string(ln) + ": bad value for " + k;
            return false;
        }
    }
    return true;
}

// Blocking FIFO; pop returns false once closed and drained.
template <class T>
class Queue {
public:
    void push(T v) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.push_back(std::move(v));
        }
        cv_.notify_one();
    }
    bool pop(T& v) {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return !q_.empty() || closed_; });
        if (q_.empty()) return false;
        v = std::move(q_.front());
        q_.pop_front();
        return true;
    }
    void close() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            closed_ = true;
        }
        cv_.notify_all();
    }

private:
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<T> q_;
    bool closed_ = false;
};

struct KeyItem {
    int buf;
    uint32_t m, flags;
};
struct SeedItem {
    int buf;
    int32_t reuse;
};
struct Done {
    pa_ticket ticket = 0;
    int key_buf = -1, out_buf = -1, seed_buf = -1;  // buffers to recycle once the block is out
    pa_seed_id drop_seed = 0;                         // resident seed no later block uses
    uint32_t n = 0, m = 0, flags = 0;
    bool last = false;
};

std::mutex g_print;
#define say(...)                                          \
    do {                                                  \
        std::lock_guard<std::mutex> lk_(g_print);         \
        std::printf(__VA_ARGS__);                         \
        std::fflush(stdout);                              \
    } while (0)

bool read_file(const std::string& path, void* dst, size_t bytes, int missing_code, int short_code, int& code) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        say("Can't open file \"%s\" => terminating!\n", path.c_str());
        code = missing_code;
        return false;
    }
    f.read(static_cast<char*>(dst), std::streamsize(bytes));
    if (size_t(f.gcount()) != bytes) {
        say("File \"%s\" is too short: %zu bytes required => terminating!\n", path.c_str(), bytes);
        code = short_code;
        return false;
    }
    return true;
}

bool parse_hex32(const std::string& h, uint8_t out[32]) {
    if (h.size() != 64) return false;
    for (int i = 0; i < 32; ++i) {
        unsigned v;
        if (std::sscanf(h.c_str() + 2 * i, "%2x", &v) != 1) return false;
        out[i] = uint8_t(v);
    }
    return true;
}

// argv "speedtest" (main, PrivacyAmplification.cu:1631-1643; timing :1758-1777, 2077-2089): with a
// fresh seed per block (i = 0, reuse 0) and with a resident one (i = 1, reuse -1), for every
// n = 2^j, j = 10..27, m = 3n/8: 4 untimed blocks, then 10 timed ones, each printed as
// d[i,j,k]=<Mbit/s of input key> (n over the block's wall time, host buffers in and out).
int run_speedtest(pa_engine* eng) {
    constexpr uint32_t JMAX = 27;
    const uint64_t nmax = 1ull << JMAX;
    uint32_t *key = nullptr, *seed = nullptr;
    uint8_t* out = nullptr;
    int rc = pa_host_alloc(eng, (nmax / 32 + 2) * 4, reinterpret_cast<void**>(&key));
    if (rc == PA_OK) rc = pa_host_alloc(eng, nmax / 8, reinterpret_cast<void**>(&seed));
    if (rc == PA_OK) rc = pa_host_alloc(eng, nmax / 16, reinterpret_cast<void**>(&out));
    for (int i = 0; i < 2 && rc == PA_OK; ++i) {
        for (uint32_t j = 10; j <= JMAX && rc == PA_OK; ++j) {
            const uint32_t n = 1u << j, m = n / 4 + n / 8;
            pa_synth_bits(0x1000 + j, 0.5, uint64_t(n) + 1, key);
            pa_synth_bits(0x2000 + j, 0.5, n, seed);
            pa_seed_id sid = 0;
            if (i == 1 && (rc = pa_seed_put(eng, seed, n, m, -1, &sid)) != PA_OK) break;
            for (int k = -4; k < 10 && rc == PA_OK; ++k) {
                const auto t0 = std::chrono::steady_clock::now();
                rc = i ? pa_run_block_seeded(eng, key, sid, n, m, PA_XOR_KEY_REST | PA_COMPRESS, out)
                       : pa_run_block(eng, key, seed, n, m, PA_XOR_KEY_REST | PA_COMPRESS, out);
                const double us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
                if (rc == PA_OK && k >= 0) say("d[%d,%u,%d]=%.3f\n", i, j, k, n / us);
            }
            if (sid) pa_seed_drop(eng, sid);
        }
    }
    if (rc != PA_OK) say("speedtest failed: %s\n", pa_strerror(rc));
    for (void* p : {static_cast<void*>(key), static_cast<void*>(seed), static_cast<void*>(out)})
        if (p) pa_host_free(eng, p);
    return rc == PA_OK ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    std::string cfg_path = "config.yaml";
    int64_t max_blocks = -1;
    int factor_override = -1;
    bool speedtest = false;
    std::string unit_test;  // argv unitTest<Name> (PrivacyAmplification.cu:1645-1652)
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--config" && i + 1 < argc) cfg_path = argv[++i];
        else if (a == "--factor_exp" && i + 1 < argc) factor_override = std::atoi(argv[++i]);
        else if (a == "--blocks" && i + 1 < argc) max_blocks = std::atoll(argv[++i]);
        else if (a == "speedtest") speedtest = true;
        else if (a.compare(0, 8, "unitTest") == 0) unit_test = a.substr(8);
    }
    Config cfg;
    std::string err;
    if (!load_config(cfg_path, cfg, err)) {
        say("%s\nCan't read the configuration => terminating!\n", err.c_str());
        return 102;
    }
    if (factor_override >= 0) cfg.factor_exp = uint32_t(factor_override);
    if (cfg.factor_exp < 10 || cfg.factor_exp > 29) {
        say("factor_exp %u is outside 10..29 => terminating!\n", cfg.factor_exp);
        return 400;
    }
    const uint32_t n = 1u << cfg.factor_exp;
    const uint32_t nw = n / 32;
    const uint32_t ncache = std::max(2u, cfg.input_blocks_to_cache), ocache = std::max(2u, cfg.output_blocks_to_cache);
    uint8_t expect[32] = {};
    if (cfg.verify_ampout && !parse_hex32(cfg.verify_ampout_sha3, expect)) {
        say("verify_ampout_sha3 is not 64 hex digits => terminating!\n");
        return 102;
    }

    pa_config pc;
    pa_config_default(&pc);
    pc.device = int32_t(cfg.gpu_device_id_to_use);
    pc.max_n_log2 = std::max(27u, cfg.factor_exp);
    pc.reduction_exp = cfg.reduction_exp;
    pc.pre_mul_reduction_exp = cfg.pre_mul_reduction_exp;
    pa_engine* eng = nullptr;
    int rc = pa_create(&pc, &eng);
    if (rc != PA_OK) {
        say("Engine start failed: %s => terminating!\n", pa_strerror(rc));
        return 1;
    }
    if (speedtest) {
        const int code = run_speedtest(eng);
        pa_destroy(eng);
        return code;
    }
    if (!unit_test.empty()) {  // the engine's self-test of that step; 100 on failure, as the reference
        say("Started %s Unit Test...\n", unit_test.c_str());
        char report[512];
        const int st = pa_selftest(eng, unit_test.c_str(), report, sizeof(report));
        pa_destroy(eng);
        if (st == PA_ERR_ARG) {
            say("Unknown unit test unitTest%s\n", unit_test.c_str());
            return 1;
        }
        say("%s\n", st == PA_OK ? report : (st == PA_ERR_CHECK ? report : pa_strerror(st)));
        say("Completed %s Unit Test\n", unit_test.c_str());
        return st == PA_OK ? 0 : 100;
    }
    pa_serve_config sc{};
    sc.slots = 3;
    sc.max_batch = 64;
    sc.hbm_budget_bytes = 16ull << 30;
    if ((rc = pa_serve_start(eng, &sc)) != PA_OK) {
        say("Engine start failed: %s => terminating!\n", pa_strerror(rc));
        return 1;
    }
    // pinned block caches (the reference's cudaMallocHost'd caches, PrivacyAmplification.cu:1535-1551)
    std::vector<uint32_t*> keys(ncache), seeds(ncache);
    std::vector<uint8_t*> outs(ocache);
    for (auto& p : keys) rc |= pa_host_alloc(eng, (nw + 2) * 4ull, reinterpret_cast<void**>(&p));
    for (auto& p : seeds) rc |= pa_host_alloc(eng, nw * 4ull, reinterpret_cast<void**>(&p));
    // m/8 bytes for any vertical_block the key socket accepts (m < n), n/16 without compression
    for (auto& p : outs) rc |= pa_host_alloc(eng, n / 8 + 16, reinterpret_cast<void**>(&p));
    if (rc != PA_OK) {
        say("Pinned buffer allocation failed => terminating!\n");
        return 1;
    }
    Queue<int> free_keys, free_seeds, free_outs;
    for (uint32_t i = 0; i < ncache; ++i) {
        free_keys.push(int(i));
        free_seeds.push(int(i));
    }
    for (uint32_t i = 0; i < ocache; ++i) free_outs.push(int(i));
    Queue<KeyItem> key_q;
    Queue<SeedItem> seed_q;
    Queue<Done> done_q;
    int fatal = 0;
    std::mutex fatal_mu;
    // A fatal error anywhere closes the input queues, so the block loop's waits on them return
    // and the process ends with the error's code instead of waiting for more input.
    auto set_fatal = [&](int code) {
        {
            std::lock_guard<std::mutex> lk(fatal_mu);
            if (!fatal) fatal = code;
        }
        key_q.close();
        seed_q.close();
    };

    // ---- inputs
    zmtp::Endpoint* key_ep = cfg.use_key_server ? new zmtp::Endpoint("PULL", cfg.address_key_in, false) : nullptr;
    zmtp::Endpoint* seed_ep = cfg.use_matrix_seed_server ? new zmtp::Endpoint("PULL", cfg.address_seed_in, false) : nullptr;
    std::vector<uint32_t> file_key, file_seed;
    if (!cfg.use_key_server) {
        file_key.assign(nw + 2, 0u);
        int code = 0;
        if (!read_file(cfg.keyfile_path, file_key.data(), (nw + 1) * 4ull, 105, 106, code)) return code;
    }
    if (!cfg.use_matrix_seed_server) {
        file_seed.assign(nw, 0u);
        int code = 0;
        if (!read_file(cfg.toeplitz_seed_path, file_seed.data(), nw * 4ull, 103, 104, code)) return code;
    }
    std::thread key_rx([&] {
        struct Closer { Queue<KeyItem>& q; ~Closer() { q.close(); } } closer{key_q};  // no more keys
        while (true) {
            int b;
            if (!free_keys.pop(b)) return;
            KeyItem it{b, uint32_t(cfg.vertical_len), (cfg.do_xor_key_rest ? uint32_t(PA_XOR_KEY_REST) : 0u) | (cfg.do_compress ? uint32_t(PA_COMPRESS) : 0u)};
            if (!key_ep) {
                std::memcpy(keys[b], file_key.data(), (nw + 2) * 4ull);
                key_q.push(it);
                continue;
            }
            zmtp::Message msg;
            while (true) {
                if (!key_ep->recv(msg)) return;
                // [u8 do_xor_key_rest][u8 do_compress][u32 vertical_block][key words]
                if (msg.size() == 4 && msg[0].size() == 1 && msg[1].size() == 1 && msg[2].size() == 4) {
                    uint32_t vb;
                    std::memcpy(&vb, msg[2].data(), 4);
                    const bool x = msg[0][0] != 0, c = msg[1][0] != 0;
                    const uint64_t want = x ? (nw + 1) * 4ull : ((n - 32ull * vb) / 32 + 1) * 4ull;
                    if (vb > 0 && 32ull * vb < n && msg[3].size() == want) {
                        std::memset(keys[b], 0, (nw + 2) * 4ull);
                        std::memcpy(keys[b], msg[3].data(), msg[3].size());
                        it.m = 32u * vb;
                        it.flags = (x ? uint32_t(PA_XOR_KEY_REST) : 0u) | (c ? uint32_t(PA_COMPRESS) : 0u);
                        break;
                    }
                }
                say("Error receiving data from Keyserver! Retrying...\n");
            }
            if (cfg.show_zeromq_status) say("Key Block recived\n");
            key_q.push(it);
        }
    });
    std::thread seed_rx([&] {
        struct Closer { Queue<SeedItem>& q; ~Closer() { q.close(); } } closer{seed_q};  // no more seeds
        while (true) {
            int b;
            if (!free_seeds.pop(b)) return;
            SeedItem it{b, cfg.reuse_seed_amount};
            if (!seed_ep) {
                std::memcpy(seeds[b], file_seed.data(), nw * 4ull);
                seed_q.push(it);
                continue;
            }
            zmtp::Message msg;
            while (true) {
                if (!seed_ep->recv(msg)) return;
                // [int32 reuse_seed_amount][n/8 bytes seed]
                if (msg.size() == 2 && msg[0].size() == 4 && msg[1].size() == nw * 4ull) {
                    std::memcpy(&it.reuse, msg[0].data(), 4);
                    std::memcpy(seeds[b], msg[1].data(), nw * 4ull);
                    break;
                }
                say("Error receiving data from Seedserver! Retrying...\n");
            }
            if (cfg.show_zeromq_status) say("Seed Block recived\n");
            seed_q.push(it);
        }
    });

    // ---- output verification: SHA3-256 of each output on verify_ampout_threads workers (the
    // reference's verifyDataPool, PrivacyAmplification.cu:1160-1180, 1130-1140), off the send path.
    // An output buffer returns to the free list once its worker has hashed it.
    struct VerifyJob {
        int out_buf;
        size_t bytes;
    };
    Queue<VerifyJob> verify_q;
    std::vector<std::thread> verifiers;
    if (cfg.verify_ampout)
        for (uint32_t i = 0; i < std::max(1u, cfg.verify_ampout_threads); ++i)
            verifiers.emplace_back([&] {
                VerifyJob j;
                while (verify_q.pop(j)) {
                    uint8_t dg[32];
                    pa_sha3::sha3_256(outs[j.out_buf], j.bytes, dg);
                    if (std::memcmp(dg, expect, 32) != 0) {
                        say("VERIFICATION FAILED!\n");
                        set_fatal(101);
                    } else {
                        say("VERIFIED!\n");
                    }
                    free_outs.push(j.out_buf);
                }
            });

    // ---- outputs, in block order
    zmtp::Endpoint* out_ep = cfg.host_ampout_server ? new zmtp::Endpoint("PUSH", cfg.address_amp_out, true) : nullptr;
    std::thread tx([&] {
        std::ofstream ampout;
        int32_t to_store = cfg.store_first_ampouts_in_file;
        if (to_store != 0) ampout.open("ampout.bin", std::ios::binary);
        auto t_prev = std::chrono::steady_clock::now();
        Done d;
        while (done_q.pop(d)) {
            {
                std::lock_guard<std::mutex> lk(fatal_mu);
                if (fatal) break;
            }
            if (d.ticket) {
                uint64_t lat = 0;
                const int st = pa_serve_wait(eng, d.ticket, -1, &lat);
                if (st != PA_OK) {
                    say("Block failed: %s => terminating!\n", pa_strerror(st));
                    set_fatal(1);
                    break;
                }
                const size_t bytes = (d.flags & PA_COMPRESS) ? d.m / 8 : d.n / 16;
                const uint8_t* o = outs[d.out_buf];
                if (out_ep && !out_ep->send(zmtp::Message{zmtp::Frame(o, o + bytes)})) {
                    set_fatal(1);
                    break;
                }
                if (out_ep && cfg.show_zeromq_status) say("Block sent to AMPOUT Client\n");
                if (to_store != 0 && ampout) {
                    ampout.write(reinterpret_cast<const char*>(o), std::streamsize(bytes));
                    ampout.flush();
                    if (to_store > 0 && --to_store == 0) ampout.close();
                }
                const auto now = std::chrono::steady_clock::now();
                const double ms = std::chrono::duration<double, std::milli>(now - t_prev).count();
                t_prev = now;
                if (cfg.show_ampout >= 0) {
                    say("Blocktime: %.3f ms => %.1f Mbit/s\n", ms, d.n / (ms * 1e3));
                    for (int i = 0; i < cfg.show_ampout && size_t(i) < bytes; ++i) {  // byte and its bits, MSB first
                        char bin[9] = {};
                        for (int k = 0; k < 8; ++k) bin[k] = char('0' + ((o[i] >> (7 - k)) & 1));
                        say("0x%02X: %s\n", o[i], bin);
                    }
                }
                if (cfg.verify_ampout) verify_q.push(VerifyJob{d.out_buf, bytes});  // the worker frees it
                else free_outs.push(d.out_buf);
                free_keys.push(d.key_buf);
            }
            if (d.seed_buf >= 0) free_seeds.push(d.seed_buf);
            if (d.drop_seed) pa_seed_drop(eng, d.drop_seed);
            if (d.last) break;
        }
        verify_q.close();
        for (auto& t : verifiers) t.join();
        free_keys.close();
        free_seeds.close();
        free_outs.close();
        key_q.close();
        seed_q.close();
    });

    // ---- blocks (the mainloop): seed choice, submission
    int32_t reuse = 0;
    bool first = true;
    int dyn_seed_buf = -1;        // current seed held for dynamic blocks
    pa_seed_id resident = 0;      // current resident seed
    for (int64_t blk = 0; max_blocks < 0 || blk < max_blocks; ++blk) {
        {
            std::lock_guard<std::mutex> lk(fatal_mu);
            if (fatal) break;
        }
        KeyItem k;
        if (!key_q.pop(k)) break;
        Done d;
        if (reuse == 0 || first) {  // take the next seed (mainloop, PrivacyAmplification.cu:1781-1791)
            SeedItem s;
            if (!seed_q.pop(s)) break;
            d.seed_buf = dyn_seed_buf;  // the previous seed's buffer is free once the earlier blocks are out
            d.drop_seed = resident;
            dyn_seed_buf = -1;
            resident = 0;
            reuse = s.reuse;
            first = false;
            if (reuse != 0 && (n & (n - 1)) == 0) {
                rc = pa_seed_put(eng, seeds[s.buf], n, k.m, reuse, &resident);
                if (rc != PA_OK) {
                    say("Seed upload failed: %s => terminating!\n", pa_strerror(rc));
                    set_fatal(1);
                    break;
                }
                free_seeds.push(s.buf);
            } else {
                dyn_seed_buf = s.buf;
            }
        }
        int ob;
        if (!free_outs.pop(ob)) break;
        pa_block b{};
        b.key = keys[k.buf];
        b.seed = resident ? nullptr : seeds[dyn_seed_buf];
        b.seed_id = resident;
        b.n = n;
        b.m = k.m;
        b.flags = k.flags;
        b.out = outs[ob];
        rc = pa_serve_submit(eng, &b, &d.ticket);
        if (rc != PA_OK) {
            say("Block submission failed: %s => terminating!\n", pa_strerror(rc));
            set_fatal(1);
            break;
        }
        d.key_buf = k.buf;
        d.out_buf = ob;
        d.n = n;
        d.m = k.m;
        d.flags = k.flags;
        d.last = max_blocks >= 0 && blk + 1 == max_blocks;
        done_q.push(d);
        if (reuse > 0) --reuse;  // PrivacyAmplification.cu:2099-2101
    }
    done_q.close();
    tx.join();
    if (key_ep) key_ep->stop();
    if (seed_ep) seed_ep->stop();
    free_keys.close();
    free_seeds.close();
    key_rx.join();
    seed_rx.join();
    delete key_ep;
    delete seed_ep;
    delete out_ep;
    pa_serve_stop(eng);
    pa_destroy(eng);
    return fatal;
}

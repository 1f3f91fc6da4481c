// zmtp.cpp -- ZMTP/3.0 NULL-mechanism PUSH/PULL endpoints (see zmtp.hpp).
#include "zmtp.hpp"

#include <arpa/inet.h>
#include <netdb.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <poll.h>
#include <sys/socket.h>
#include <unistd.h>

#include <chrono>
#include <cstring>
#include <thread>

namespace zmtp {

namespace {
constexpr uint8_t F_MORE = 1, F_LONG = 2, F_COMMAND = 4;
constexpr size_t MAX_FRAME = size_t(1) << 31;  // refuse absurd sizes from a corrupt stream

void put_be64(uint8_t* p, uint64_t v) {
    for (int i = 7; i >= 0; --i) {
        p[i] = uint8_t(v & 0xff);
        v >>= 8;
    }
}
uint64_t get_be64(const uint8_t* p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v = (v << 8) | p[i];
    return v;
}
}  // namespace

Endpoint::Endpoint(std::string role, std::string address, bool bind) : role_(std::move(role)), bind_(bind) {
    if (!parse_address(address, host_, port_)) {
        host_.clear();
        port_ = 0;
    }
}

Endpoint::~Endpoint() {
    drop();
    if (listen_fd_ >= 0) ::close(listen_fd_);
}

bool Endpoint::parse_address(const std::string& a, std::string& host, int& port) {
    const std::string pre = "tcp://";
    if (a.compare(0, pre.size(), pre) != 0) return false;
    const size_t colon = a.rfind(':');
    if (colon == std::string::npos || colon < pre.size()) return false;
    host = a.substr(pre.size(), colon - pre.size());
    const std::string ps = a.substr(colon + 1);
    if (host.empty() || ps.empty() || ps.size() > 5) return false;
    port = 0;
    for (char c : ps) {
        if (c < '0' || c > '9') return false;
        port = port * 10 + (c - '0');
    }
    return port > 0 && port < 65536;
}

std::vector<uint8_t> Endpoint::greeting(bool as_server) {
    std::vector<uint8_t> g(64, 0);
    g[0] = 0xff;  // signature: %xFF, 8 padding octets, %x7F
    g[8] = 0x01;
    g[9] = 0x7f;
    g[10] = 3;  // version 3.0
    g[11] = 0;
    std::memcpy(&g[12], "NULL", 4);  // mechanism, zero-padded to 20 octets
    g[32] = as_server ? 1 : 0;
    return g;
}

std::vector<uint8_t> Endpoint::ready_command(const std::string& socket_type) {
    std::vector<uint8_t> body;
    const char* name = "READY";
    body.push_back(5);
    body.insert(body.end(), name, name + 5);
    const std::string key = "Socket-Type";
    body.push_back(uint8_t(key.size()));
    body.insert(body.end(), key.begin(), key.end());
    const uint32_t vl = uint32_t(socket_type.size());
    for (int s = 24; s >= 0; s -= 8) body.push_back(uint8_t(vl >> s));
    body.insert(body.end(), socket_type.begin(), socket_type.end());
    std::vector<uint8_t> f;
    f.push_back(F_COMMAND);
    f.push_back(uint8_t(body.size()));
    f.insert(f.end(), body.begin(), body.end());
    return f;
}

void Endpoint::stop() {
    stop_ = true;
    const int fd = fd_.load();
    if (fd >= 0) ::shutdown(fd, SHUT_RDWR);  // wakes a blocked transfer; loops also poll stop_
}

void Endpoint::drop() {
    if (fd_ >= 0) {
        ::close(fd_);
        fd_ = -1;
    }
}

bool Endpoint::read_exact(void* p, size_t n) {
    uint8_t* b = static_cast<uint8_t*>(p);
    while (n) {
        if (stop_) return false;
        pollfd pf{fd_, POLLIN, 0};
        const int pr = ::poll(&pf, 1, 200);
        if (pr == 0) continue;
        if (pr < 0) return false;
        const ssize_t r = ::recv(fd_, b, n, 0);
        if (r <= 0) return false;
        b += r;
        n -= size_t(r);
    }
    return true;
}

bool Endpoint::write_all(const void* p, size_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(p);
    while (n) {
        if (stop_) return false;
        pollfd pf{fd_, POLLOUT, 0};
        const int pr = ::poll(&pf, 1, 200);
        if (pr == 0) continue;
        if (pr < 0) return false;
        const ssize_t r = ::send(fd_, b, n, MSG_NOSIGNAL);
        if (r <= 0) return false;
        b += r;
        n -= size_t(r);
    }
    return true;
}

bool Endpoint::read_frame(uint8_t& flags, Frame& body) {
    if (!read_exact(&flags, 1)) return false;
    uint64_t size;
    if (flags & F_LONG) {
        uint8_t s8[8];
        if (!read_exact(s8, 8)) return false;
        size = get_be64(s8);
    } else {
        uint8_t s1;
        if (!read_exact(&s1, 1)) return false;
        size = s1;
    }
    if (size > MAX_FRAME) return false;
    body.resize(size_t(size));
    return size == 0 || read_exact(body.data(), body.size());
}

bool Endpoint::handshake() {
    const auto mine = greeting(bind_);
    if (!write_all(mine.data(), mine.size())) return false;
    uint8_t peer[64];
    if (!read_exact(peer, sizeof(peer))) return false;
    if (peer[0] != 0xff || peer[9] != 0x7f || peer[10] < 3) return false;
    if (std::memcmp(&peer[12], "NULL", 4) != 0 || peer[16] != 0) return false;
    const auto ready = ready_command(role_);
    if (!write_all(ready.data(), ready.size())) return false;
    uint8_t flags;
    Frame cmd;
    if (!read_frame(flags, cmd) || !(flags & F_COMMAND)) return false;
    return cmd.size() >= 6 && cmd[0] == 5 && std::memcmp(&cmd[1], "READY", 5) == 0;
}

bool Endpoint::ensure() {
    while (!stop_) {
        if (fd_ >= 0) return true;
        if (port_ == 0) return false;
        if (bind_) {
            if (listen_fd_ < 0) {
                listen_fd_ = ::socket(AF_INET, SOCK_STREAM, 0);
                const int one = 1;
                ::setsockopt(listen_fd_, SOL_SOCKET, SO_REUSEADDR, &one, sizeof(one));
                sockaddr_in sa{};
                sa.sin_family = AF_INET;
                sa.sin_port = htons(uint16_t(port_));
                if (host_ == "*") sa.sin_addr.s_addr = htonl(INADDR_ANY);
                else if (inet_pton(AF_INET, host_ == "localhost" ? "127.0.0.1" : host_.c_str(), &sa.sin_addr) != 1)
                    return false;
                if (::bind(listen_fd_, reinterpret_cast<sockaddr*>(&sa), sizeof(sa)) != 0 ||
                    ::listen(listen_fd_, 1) != 0) {
                    ::close(listen_fd_);
                    listen_fd_ = -1;
                    std::this_thread::sleep_for(std::chrono::milliseconds(100));  // bind retried
                    continue;
                }
            }
            pollfd pf{listen_fd_, POLLIN, 0};
            if (::poll(&pf, 1, 200) <= 0) continue;
            fd_ = ::accept(listen_fd_, nullptr, nullptr);
        } else {
            addrinfo hints{}, *res = nullptr;
            hints.ai_family = AF_INET;
            hints.ai_socktype = SOCK_STREAM;
            const std::string port = std::to_string(port_);
            if (::getaddrinfo(host_.c_str(), port.c_str(), &hints, &res) != 0 || !res) {
                std::this_thread::sleep_for(std::chrono::milliseconds(100));
                continue;
            }
            fd_ = ::socket(res->ai_family, res->ai_socktype, res->ai_protocol);
            if (fd_ >= 0 && ::connect(fd_, res->ai_addr, res->ai_addrlen) != 0) drop();
            ::freeaddrinfo(res);
            if (fd_ < 0) {
                std::this_thread::sleep_for(std::chrono::milliseconds(100));  // connect retried
                continue;
            }
        }
        if (fd_ < 0) continue;
        const int one = 1;
        ::setsockopt(fd_, IPPROTO_TCP, TCP_NODELAY, &one, sizeof(one));
        if (!handshake()) drop();
    }
    return false;
}

bool Endpoint::recv(Message& msg) {
    while (ensure()) {
        msg.clear();
        bool ok = true;
        while (true) {
            uint8_t flags;
            Frame body;
            if (!read_frame(flags, body)) {
                ok = false;
                break;
            }
            if (flags & F_COMMAND) continue;  // heartbeats and the like carry no data
            msg.push_back(std::move(body));
            if (!(flags & F_MORE)) break;
        }
        if (ok) return true;
        drop();  // retry on a fresh connection
    }
    return false;
}

bool Endpoint::send(const Message& msg) {
    while (ensure()) {
        bool ok = true;
        for (size_t i = 0; i < msg.size() && ok; ++i) {
            const Frame& f = msg[i];
            uint8_t hdr[9];
            size_t hl;
            const uint8_t more = i + 1 < msg.size() ? F_MORE : 0;
            if (f.size() > 255) {
                hdr[0] = uint8_t(more | F_LONG);
                put_be64(hdr + 1, f.size());
                hl = 9;
            } else {
                hdr[0] = more;
                hdr[1] = uint8_t(f.size());
                hl = 2;
            }
            ok = write_all(hdr, hl) && (f.empty() || write_all(f.data(), f.size()));
        }
        if (ok) return true;
        drop();
    }
    return false;
}

}  // namespace zmtp

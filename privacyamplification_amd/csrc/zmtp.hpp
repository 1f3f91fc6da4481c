// zmtp.hpp -- minimal ZMTP/3.0 (NULL security) PUSH/PULL endpoints over TCP.
//
// The reference talks to its key server, seed server and amp-out client through ZeroMQ PULL/PUSH
// sockets with multipart messages (PrivacyAmplification.cu:980-1073 receive, 1143-1200 send;
// peers in examples/*).  This is the wire protocol those libzmq peers speak (ZMTP 3.0, RFC 23):
// a 64-byte greeting, a NULL-mechanism READY command carrying Socket-Type, then frames
//   flags (bit0 MORE, bit1 LONG, bit2 COMMAND) | size (1 or 8 octets, big-endian) | body.
// One peer per endpoint; a lost connection is re-established and the message retried, as the
// reference retries failed receives and sends.
#pragma once
#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

namespace zmtp {

using Frame = std::vector<uint8_t>;
using Message = std::vector<Frame>;

class Endpoint {
public:
    // role: "PUSH" or "PULL"; address: tcp://host:port.  bind == true listens for one peer.
    Endpoint(std::string role, std::string address, bool bind);
    ~Endpoint();
    Endpoint(const Endpoint&) = delete;
    Endpoint& operator=(const Endpoint&) = delete;

    // Blocks until a whole message arrives (reconnecting as needed); false once stop() was called.
    bool recv(Message& msg);
    // Blocks until the message is written (reconnecting as needed); false once stop() was called.
    bool send(const Message& msg);
    // Makes blocked and later calls return false (from any thread).
    void stop();

    // Greeting and READY command this endpoint sends (exposed for tests).
    static std::vector<uint8_t> greeting(bool as_server);
    static std::vector<uint8_t> ready_command(const std::string& socket_type);
    // Parses "tcp://host:port"; false on a malformed address.
    static bool parse_address(const std::string& a, std::string& host, int& port);

private:
    bool ensure();              // connected and handshaken
    bool handshake();
    bool read_exact(void* p, size_t n);
    bool write_all(const void* p, size_t n);
    bool read_frame(uint8_t& flags, Frame& body);
    void drop();

    std::string role_, host_;
    int port_ = 0;
    bool bind_ = false;
    int listen_fd_ = -1;
    std::atomic<int> fd_{-1};  // read by stop() from another thread
    std::atomic<bool> stop_{false};
};

}  // namespace zmtp

// zmtp_c.cpp -- C ABI of libpa_zmtp.so (include/pa_zmtp.h).
#include "../../include/pa_zmtp.h"

#include <memory>

#include "sha3.hpp"
#include "zmtp.hpp"

struct pa_zmtp {
    std::unique_ptr<zmtp::Endpoint> ep;
    zmtp::Message last;
};

extern "C" {

pa_zmtp* pa_zmtp_open(const char* role, const char* address, int bind) {
    if (!role || !address) return nullptr;
    const std::string r(role);
    if (r != "PUSH" && r != "PULL") return nullptr;
    std::string host;
    int port;
    if (!zmtp::Endpoint::parse_address(address, host, port)) return nullptr;
    auto z = new pa_zmtp();
    z->ep.reset(new zmtp::Endpoint(r, address, bind != 0));
    return z;
}

int pa_zmtp_send(pa_zmtp* z, const void* const* parts, const uint64_t* lens, uint32_t nparts) {
    if (!z || (nparts && (!parts || !lens))) return -1;
    zmtp::Message m(nparts);
    for (uint32_t i = 0; i < nparts; ++i) {
        const uint8_t* p = static_cast<const uint8_t*>(parts[i]);
        m[i].assign(p, p + lens[i]);
    }
    return z->ep->send(m) ? 0 : -1;
}

int pa_zmtp_recv(pa_zmtp* z, uint32_t* nparts) {
    if (!z || !nparts) return -1;
    if (!z->ep->recv(z->last)) return -1;
    *nparts = uint32_t(z->last.size());
    return 0;
}

int pa_zmtp_part(pa_zmtp* z, uint32_t i, const void** data, uint64_t* len) {
    if (!z || !data || !len || i >= z->last.size()) return -1;
    *data = z->last[i].data();
    *len = z->last[i].size();
    return 0;
}

void pa_zmtp_stop(pa_zmtp* z) {
    if (z) z->ep->stop();
}

void pa_zmtp_close(pa_zmtp* z) { delete z; }

void pa_sha3_256(const void* data, uint64_t len, uint8_t digest[32]) {
    pa_sha3::sha3_256(static_cast<const uint8_t*>(data), size_t(len), digest);
}

}  // extern "C"

// sha3.hpp -- SHA3-256 (FIPS 202: Keccak-f[1600], rate 136 bytes, domain suffix 0x06).
// Used to verify outputs against an expected digest, as the reference's verify_ampout does
// (PrivacyAmplification.cu:1120-1140).
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace pa_sha3 {

inline uint64_t rotl(uint64_t x, int n) { return n ? (x << n) | (x >> (64 - n)) : x; }

inline void keccak_f1600(uint64_t a[25]) {
    static const uint64_t RC[24] = {
        0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
        0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
        0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
        0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
        0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
        0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};
    // rho offsets r[x][y] and pi: lane (x, y) moves to (y, 2x + 3y)
    static const int R[5][5] = {{0, 36, 3, 41, 18}, {1, 44, 10, 45, 2}, {62, 6, 43, 15, 61},
                                {28, 55, 25, 21, 56}, {27, 20, 39, 8, 14}};
    for (int round = 0; round < 24; ++round) {
        uint64_t c[5], d[5], b[25];
        for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
        for (int x = 0; x < 5; ++x) d[x] = c[(x + 4) % 5] ^ rotl(c[(x + 1) % 5], 1);
        for (int i = 0; i < 25; ++i) a[i] ^= d[i % 5];
        for (int x = 0; x < 5; ++x)
            for (int y = 0; y < 5; ++y) b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl(a[x + 5 * y], R[x][y]);
        for (int x = 0; x < 5; ++x)
            for (int y = 0; y < 5; ++y)
                a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
        a[0] ^= RC[round];
    }
}

// digest = SHA3-256(data[0, len))
inline void sha3_256(const uint8_t* data, size_t len, uint8_t digest[32]) {
    constexpr size_t RATE = 136;
    uint64_t a[25] = {};
    auto absorb = [&](const uint8_t* blk) {
        for (size_t i = 0; i < RATE / 8; ++i) {
            uint64_t lane = 0;
            for (int k = 7; k >= 0; --k) lane = (lane << 8) | blk[8 * i + size_t(k)];
            a[i] ^= lane;
        }
        keccak_f1600(a);
    };
    while (len >= RATE) {
        absorb(data);
        data += RATE;
        len -= RATE;
    }
    uint8_t last[RATE] = {};
    if (len) std::memcpy(last, data, len);
    last[len] ^= 0x06;
    last[RATE - 1] ^= 0x80;
    absorb(last);
    for (int i = 0; i < 32; ++i) digest[i] = uint8_t(a[i / 8] >> (8 * (i % 8)));
}

}  // namespace pa_sha3

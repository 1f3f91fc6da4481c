// pa_geom.hpp -- block geometry shared by the host planner and the kernels.
//
// A block (n, m, flags) is evaluated as a circular convolution of length L (a power of two, L >= n)
// of the key x (kbits = h+1 bits, h = n-m) with the seed t (n bits, rotated left by m when L > n),
// read at conv indices [out_off, out_off + out_bits).  For power-of-two n, L = n and out_off = 0,
// which is the reference's own circular form y[p] = XOR_j x[j] s[(p-j) mod n].
//
// The length-L real transform is a length-N complex transform (N = L/2, z[t] = x[2t] + i x[2t+1])
// evaluated as a four-step FFT with N = N1 * N2 (t = t2 + N2*t1, k = k1 + N1*k2):
//   pass A  (fwd_cols)  column FFTs of length N1, decimated by D classes of k1 (M = N1/D per class),
//                       16 columns per workgroup, input from the bit-transposed key/seed;
//   pass B  (rows)      row FFTs of length N2 on rows k1 and N1-k1, real split, DC zeroing, complex
//                       product, inverse split, inverse row FFTs;
//   pass A' (inv_cols)  inverse column FFTs of length N1, W columns per workgroup, round-mod-2 epilogue
//                       (N1 = 2^15: one column per workgroup, 2-bit crumbs instead of nibbles).
// N <= 8192 (L <= 2^14) runs fully on chip in one workgroup (small path).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pa {

constexpr int ROW_LOG = 13;            // N2 = 8192 for the four-step path
constexpr int SMALL_MAX_LOG_N = 13;    // N <= 8192 -> single-workgroup path
constexpr int MAX_L_LOG = 29;          // largest transform served (L = 2^29; the reference stops at 2^27)

// Pass B's output R'[k1][t2] when 2^RT_LG1_MIN <= N1 <= 2^RT_LG1_MAX (transforms of 2^24 to 2^29
// bits) goes to its own buffer in tiles: row pair p (rows p and N1-p, or 0 and N1/2 for p = 0), side
// s (0: row p, 1: its partner) at Rt[((t2 / RT_COLS) N1/2 + p) 2 RT_COLS + s RT_COLS + t2 mod RT_COLS].
// A tile holds both rows' RT_COLS columns (2 x 128 bytes): pass B writes whole 128-byte lines, and a
// workgroup of pass A' (1 to 8 columns) reads its columns from consecutive lines along p instead of
// from rows 8 N2 bytes apart.  At N1 = 2^RT_LG1 pass A' runs two 256-thread workgroups per CU.
constexpr int RT_LG1_MIN = 10;
constexpr int RT_LG1 = 13;
constexpr int RT_LG1_MAX = 15;
constexpr int RT_COLS = 16;
__host__ __device__ constexpr bool rt_tiled(uint32_t lg1) {
    return lg1 >= (uint32_t)RT_LG1_MIN && lg1 <= (uint32_t)RT_LG1_MAX;
}

struct Geom {
    uint32_t n, m, h, kbits;           // block shape; kbits = h + 1
    uint32_t L, N, lgN;                // transform length, half length, log2(N)
    uint32_t N1, N2, lg1, lg2;         // four-step split (N1 = 1 on the small path)
    uint32_t M, D, W;                  // pass A class length / classes, pass A' columns per workgroup
    uint32_t out_bits, out_off;        // output window in conv indices
    uint32_t seed_words, seed_rot;     // n/32, rotation of the seed in words (m/32 when L > n); the
                                       // rotated seed repeats with period n up to L
    uint32_t key_hw;                   // h/32: index of the key word masked to its MSB
    uint32_t do_xor, reverse_endian;
    float inv_N;                       // 1/(8N): output normalisation (r2c_split / c2r_split scale)
};

// One block of a batched launch: its geometry and device buffers.  The blocks of one launch share
// the transform length L (and so N1, N2, M, D, W); n, m, flags and the buffers are per block.
struct BlockArgs {
    Geom g;
    const uint32_t* key;
    const uint32_t* seed;   // dynamic seed, or nullptr for a resident one
    const float2* sspec;    // resident split spectrum (N+1 values) when seed == nullptr
    const uint32_t* ws;     // resident seed popcount partials when seed == nullptr
    float2* Yk;             // N values (four-step path only)
    float2* Ys;             // N+1 values (dynamic seed, four-step path only)
    uint32_t* tbk;
    uint32_t* tbs;
    float2* Rt;             // R' tiles (rt_tiled(lg1)), else nullptr
    uint8_t* nib;
    uint32_t* counts;       // key partial popcounts, then seed partial popcounts
    uint32_t* out;
    float* margin;          // may be nullptr
};

}  // namespace pa

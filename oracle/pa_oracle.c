/*
 * pa_oracle.c -- CPU oracle for the Toeplitz-hashing path of nicoboss/PrivacyAmplification.
 *
 * TEST INFRASTRUCTURE ONLY: linked by nothing in the product.  tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it through ctypes as the checker (and, for the baseline,
 * as the timed CPU restatement of the reference's FP32 method).
 *
 * Contents and the reference code each part restates:
 *   pa_or_synth_bits           synthetic inputs (xoshiro256**, splitmix64 seeding)
 *   pa_or_key_split            key2StartRest            PrivacyAmplification.cu:861-892
 *   pa_or_bin_int2float        binInt2float             PrivacyAmplification.cu:570-602
 *   pa_or_correction           calculateCorrectionFloat PrivacyAmplification.cu:363-373
 *   pa_or_element_wise_product ElementWiseProduct       PrivacyAmplification.cu:465-481
 *   pa_or_to_binary_array      ToBinaryArray[NoXOR]     PrivacyAmplification.cu:739-813
 *   conv_direct                the GF(2) Toeplitz product the whole pipeline computes
 *                              (y[p] = XOR_j x[j] s[(p-j) mod n]; mainloop PrivacyAmplification.cu:1819-2042)
 *   conv_ntt                   the same as an exact integer convolution (NTT mod 15*2^27+1)
 *   fft_pipeline               the reference's FP32 method: binInt2float -> R2C -> setFirstElementToZero
 *                              -> ElementWiseProduct -> C2R -> ToBinaryArray
 *                              (PrivacyAmplification.cu:1831-2041); cuFFT itself is closed, so the
 *                              FFT here is an own radix-2 transform with FP64-generated twiddles.
 *   pa_or_amp_out              one block with the wire semantics of reciveDataKey/sendData
 *                              (PrivacyAmplification.cu:1030-1104, 1143-1225).
 */
#include "pa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;
void pa_or_set_threads(int nthreads) {
    g_threads = nthreads;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
}
int pa_or_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

static inline uint32_t refbit(const uint32_t* w, uint64_t p) {
    return (w[p >> 5] >> (31u - (uint32_t)(p & 31u))) & 1u;
}
static inline uint32_t bitrev32(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
}
static inline uint32_t bswap32(uint32_t x) {
    return ((x & 0xff000000u) >> 24) | ((x & 0x00ff0000u) >> 8) | ((x & 0x0000ff00u) << 8) |
           ((x & 0x000000ffu) << 24);
}

/* ---------------------------------------------------------------- synthetic bits ---- */
static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static inline uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
typedef struct { uint64_t s[4]; } xoshiro_t;
static void xo_seed(xoshiro_t* x, uint64_t seed) {
    uint64_t sm = seed;
    for (int i = 0; i < 4; ++i) x->s[i] = splitmix64(&sm);
}
static inline uint64_t xo_next(xoshiro_t* x) {
    uint64_t* s = x->s;
    uint64_t r = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return r;
}
void pa_or_synth_bits(uint64_t seed, double p, uint64_t nbits, uint32_t* out) {
    xoshiro_t x;
    xo_seed(&x, seed);
    uint64_t nwords = (nbits + 31) / 32;
    if (p == 0.5) {
        for (uint64_t i = 0; i < nwords; i += 2) {
            uint64_t r = xo_next(&x);
            out[i] = (uint32_t)(r >> 32);
            if (i + 1 < nwords) out[i + 1] = (uint32_t)r;
        }
    } else {
        memset(out, 0, nwords * 4);
        for (uint64_t b = 0; b < nbits; ++b) {
            double u = (double)(xo_next(&x) >> 11) * 0x1.0p-53;
            if (u < p) out[b >> 5] |= 1u << (31 - (b & 31));
        }
    }
    if (nbits & 31) out[nwords - 1] &= ~((1u << (32 - (nbits & 31))) - 1u);
}

/* ---------------------------------------------------------------- reference kernels ---- */
void pa_or_key_split(const uint32_t* key, uint32_t n, uint32_t m, uint32_t* key_start,
                     uint32_t* key_rest) {
    /* PrivacyAmplification.cu:861-892: memcpy of horizontal_block words, the next word masked
     * to its MSB, then key_rest[i] = (key[j] << 1) | (key[j+1] >> 31) from j = horizontal_block. */
    uint32_t hb = (n - m) / 32, vb = m / 32;
    memset(key_start, 0, (size_t)(n / 32) * 4);
    memcpy(key_start, key, (size_t)hb * 4);
    key_start[hb] = key[hb] & 0x80000000u;
    for (uint32_t i = 0, j = hb; i < vb; ++i, ++j) key_rest[i] = (key[j] << 1) | (key[j + 1] >> 31);
}

uint32_t pa_or_bin_int2float(const uint32_t* bits, uint32_t nbits, uint32_t reduction_exp,
                             float* out) {
    const float one = 1.0f / (float)(1u << reduction_exp);
    uint32_t ones = 0;
    for (uint32_t i = 0; i < nbits; ++i) {
        uint32_t b = refbit(bits, i);
        out[i] = b ? one : 0.0f;
        ones += b;
    }
    return ones;
}

float pa_or_correction(uint32_t wk, uint32_t ws, uint32_t n) {
    /* The reference multiplies two uint32 values in 32-bit arithmetic before widening. */
    uint64_t count_multiplied = (uint64_t)(uint32_t)(wk * ws);
    double normalized = (double)count_multiplied / (double)n;
    return (float)fmod(normalized, 2.0);
}

void pa_or_element_wise_product(float* a, const float* b, uint32_t count, float r) {
    for (uint32_t i = 0; i < count; ++i) {
        float ax = a[2 * i] / r, ay = a[2 * i + 1] / r;
        float bx = b[2 * i] / r, by = b[2 * i + 1] / r;
        a[2 * i] = ax * bx - ay * by;
        a[2 * i + 1] = ax * by + ay * bx;
    }
}

void pa_or_to_binary_array(const float* v, uint32_t nbits, float norm, float corr,
                           const uint32_t* key_rest, int reverse_endian, uint32_t* out) {
    uint32_t nw = (nbits + 31) / 32;
    for (uint32_t w = 0; w < nw; ++w) {
        uint32_t word = 0;
        for (uint32_t i = 0; i < 32; ++i) {
            uint32_t p = w * 32 + i;
            if (p >= nbits) break;
            float x = v[p] / norm + corr;
            int r = (int)lrintf(x); /* round half to even under the default FP environment */
            uint32_t shift = reverse_endian ? ((i & ~7u) + (7u - (i & 7u))) : (31u - i);
            word |= ((uint32_t)r & 1u) << shift;
        }
        if (key_rest) word ^= reverse_endian ? bswap32(key_rest[w]) : key_rest[w];
        out[w] = word;
    }
}

/* ---------------------------------------------------------------- direct GF(2) ---- */
/* Bit-packed schoolbook: y[p] = parity(sum_j x[j] & rev[j + n - p]) with rev[i] = s[(-i) mod n]. */
static int conv_direct(const uint32_t* ks, uint32_t kbits, const uint32_t* seed, uint32_t n,
                       uint32_t nout, uint8_t* out) {
    uint64_t kw = ((uint64_t)kbits + 63) / 64;
    uint64_t rw = (2ull * n + 63) / 64 + 2;
    uint64_t* xw = (uint64_t*)calloc(kw + 1, 8);
    uint64_t* rev = (uint64_t*)calloc(rw, 8);
    if (!xw || !rev) { free(xw); free(rev); return -1; }
    for (uint64_t j = 0; j < kbits; ++j)
        if (refbit(ks, j)) xw[j >> 6] |= 1ull << (j & 63);
    for (uint64_t i = 0; i < 2ull * n; ++i) {
        uint64_t idx = (n - (i % n)) % n;
        if (refbit(seed, idx)) rev[i >> 6] |= 1ull << (i & 63);
    }
    /* 64 pre-shifted copies make each window word a single load. */
    uint64_t** sh = (uint64_t**)malloc(64 * sizeof(uint64_t*));
    int ok = sh != NULL;
    for (int b = 0; ok && b < 64; ++b) {
        sh[b] = (uint64_t*)malloc(rw * 8);
        if (!sh[b]) { ok = 0; for (int c = 0; c < b; ++c) free(sh[c]); break; }
        for (uint64_t a = 0; a + 1 < rw; ++a)
            sh[b][a] = b ? (rev[a] >> b) | (rev[a + 1] << (64 - b)) : rev[a];
        sh[b][rw - 1] = rev[rw - 1] >> b;
    }
    if (!ok) { free(sh); free(xw); free(rev); return -1; }
    memset(out, 0, (nout + 7) / 8);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < (int64_t)nout; ++p) {
        uint64_t off = (uint64_t)n - (uint64_t)p;
        const uint64_t* w = sh[off & 63] + (off >> 6);
        uint64_t acc = 0;
        for (uint64_t k = 0; k < kw; ++k) acc ^= xw[k] & w[k];
        if (__builtin_parityll(acc)) {
#pragma omp atomic
            out[p >> 3] |= (uint8_t)(0x80u >> (p & 7));
        }
    }
    for (int b = 0; b < 64; ++b) free(sh[b]);
    free(sh); free(xw); free(rev);
    return 0;
}

/* ---------------------------------------------------------------- exact NTT ---- */
/* p = 15*2^27 + 1 (prime, primitive root 31), Montgomery form with R = 2^32. */
#define NTT_P 2013265921u
static uint32_t ntt_pinv_neg; /* -p^-1 mod 2^32 */
static uint32_t ntt_r2;       /* R^2 mod p */
static inline uint32_t mredc(uint64_t t) {
    uint32_t m = (uint32_t)t * ntt_pinv_neg;
    uint64_t u = (t + (uint64_t)m * NTT_P) >> 32;
    return (uint32_t)(u >= NTT_P ? u - NTT_P : u);
}
static inline uint32_t mmul(uint32_t a, uint32_t b) { return mredc((uint64_t)a * b); }
static inline uint32_t madd(uint32_t a, uint32_t b) { uint32_t c = a + b; return c >= NTT_P ? c - NTT_P : c; }
static inline uint32_t msub(uint32_t a, uint32_t b) { return a >= b ? a - b : a + NTT_P - b; }
static inline uint32_t to_mont(uint32_t a) { return mmul(a, ntt_r2); }
static inline uint32_t from_mont(uint32_t a) { return mredc(a); }
static uint32_t mpow(uint32_t base_m, uint64_t e) {
    uint32_t r = to_mont(1);
    while (e) { if (e & 1) r = mmul(r, base_m); base_m = mmul(base_m, base_m); e >>= 1; }
    return r;
}
static void ntt_init(void) {
    uint32_t inv = 1;
    for (int i = 0; i < 5; ++i) inv *= 2u - NTT_P * inv; /* Newton: p^-1 mod 2^32 */
    ntt_pinv_neg = (uint32_t)(0u - inv);
    uint64_t r = ((uint64_t)1 << 32) % NTT_P;
    ntt_r2 = (uint32_t)((r * r) % NTT_P);
}
/* Forward DIF (natural in, bit-reversed out) or inverse DIT (bit-reversed in, natural out). */
static void ntt_core(uint32_t* a, uint32_t n, const uint32_t* tw /* w^j, j < n/2 */, int inverse) {
    if (!inverse) {
        for (uint32_t len = n; len >= 2; len >>= 1) {
            uint32_t half = len / 2, stride = n / len;
#pragma omp parallel for schedule(static)
            for (int64_t blk = 0; blk < (int64_t)(n / len); ++blk) {
                uint32_t* x = a + (uint64_t)blk * len;
                for (uint32_t j = 0; j < half; ++j) {
                    uint32_t u = x[j], v = x[j + half];
                    x[j] = madd(u, v);
                    x[j + half] = mmul(msub(u, v), tw[(uint64_t)j * stride]);
                }
            }
        }
    } else {
        for (uint32_t len = 2; len <= n; len <<= 1) {
            uint32_t half = len / 2, stride = n / len;
#pragma omp parallel for schedule(static)
            for (int64_t blk = 0; blk < (int64_t)(n / len); ++blk) {
                uint32_t* x = a + (uint64_t)blk * len;
                for (uint32_t j = 0; j < half; ++j) {
                    uint32_t u = x[j], v = mmul(x[j + half], tw[(uint64_t)j * stride]);
                    x[j] = madd(u, v);
                    x[j + half] = msub(u, v);
                }
            }
        }
    }
}
/* a <- a (*) b cyclically over Lc (a power of two, Lc <= 2^27), Montgomery form in and out. */
static int ntt_cyclic(uint32_t* a, uint32_t* b, uint32_t Lc) {
    uint32_t* tw = (uint32_t*)malloc((size_t)(Lc / 2) * 4 + 4);
    if (!tw) return -1;
    uint32_t g = to_mont(31);
    uint32_t w = mpow(g, (uint64_t)(NTT_P - 1) / Lc); /* primitive Lc-th root */
    /* tw[j] = w^j in blocks so that the table is filled in parallel */
    {
        uint32_t half = Lc / 2, chunk = 4096;
#pragma omp parallel for schedule(static)
        for (int64_t c0 = 0; c0 < (int64_t)half; c0 += chunk) {
            uint32_t x = mpow(w, (uint64_t)c0);
            for (uint32_t j = (uint32_t)c0; j < half && j < (uint32_t)c0 + chunk; ++j) { tw[j] = x; x = mmul(x, w); }
        }
    }
    ntt_core(a, Lc, tw, 0);
    ntt_core(b, Lc, tw, 0);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)Lc; ++j) a[j] = mmul(a[j], b[j]);
    /* inverse uses w^-1 */
    {
        uint32_t winv = mpow(w, (uint64_t)Lc - 1);
        uint32_t half = Lc / 2, chunk = 4096;
#pragma omp parallel for schedule(static)
        for (int64_t c0 = 0; c0 < (int64_t)half; c0 += chunk) {
            uint32_t x = mpow(winv, (uint64_t)c0);
            for (uint32_t j = (uint32_t)c0; j < half && j < (uint32_t)c0 + chunk; ++j) { tw[j] = x; x = mmul(x, winv); }
        }
    }
    ntt_core(a, Lc, tw, 1);
    uint32_t ninv = mpow(to_mont(Lc), (uint64_t)NTT_P - 2);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)Lc; ++j) a[j] = mmul(a[j], ninv);
    free(tw);
    return 0;
}

/* Lengths above 2^27: the prime 3*2^30+1 (primitive root 5) has roots of order up to 2^30.  It
 * exceeds 2^31, so this NTT works in plain 64-bit arithmetic: twiddle products by Shoup's method
 * (precomputed floor(w 2^32 / p)), the pointwise product by a 64-bit remainder. */
#define NTT2_P 3221225473u
static int ntt_force_big = 0;  /* test hook: use the 64-bit prime at every length */
void pa_or_set_ntt_big(int on) { ntt_force_big = on; }
static inline uint32_t p2_add(uint32_t a, uint32_t b) { uint64_t s = (uint64_t)a + b; return (uint32_t)(s >= NTT2_P ? s - NTT2_P : s); }
static inline uint32_t p2_sub(uint32_t a, uint32_t b) { return a >= b ? a - b : (uint32_t)((uint64_t)a + NTT2_P - b); }
static inline uint32_t p2_mul(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) % NTT2_P); }
static inline uint32_t p2_mul_shoup(uint32_t a, uint32_t w, uint32_t wq) {
    uint64_t q = ((uint64_t)a * wq) >> 32;
    uint64_t r = (uint64_t)a * w - q * NTT2_P;
    return (uint32_t)(r >= NTT2_P ? r - NTT2_P : r);
}
static uint32_t p2_pow(uint32_t b, uint64_t e) {
    uint32_t r = 1;
    while (e) { if (e & 1) r = p2_mul(r, b); b = p2_mul(b, b); e >>= 1; }
    return r;
}
static void p2_tables(uint32_t w, uint32_t half, uint32_t* tw, uint32_t* twq) {
    uint32_t chunk = 4096;
#pragma omp parallel for schedule(static)
    for (int64_t c0 = 0; c0 < (int64_t)half; c0 += chunk) {
        uint32_t x = p2_pow(w, (uint64_t)c0);
        for (uint32_t j = (uint32_t)c0; j < half && j < (uint32_t)c0 + chunk; ++j) {
            tw[j] = x;
            twq[j] = (uint32_t)(((uint64_t)x << 32) / NTT2_P);
            x = p2_mul(x, w);
        }
    }
}
static void p2_core(uint32_t* a, uint32_t n, const uint32_t* tw, const uint32_t* twq, int inverse) {
    if (!inverse) {
        for (uint32_t len = n; len >= 2; len >>= 1) {
            uint32_t half = len / 2, stride = n / len;
#pragma omp parallel for schedule(static)
            for (int64_t blk = 0; blk < (int64_t)(n / len); ++blk) {
                uint32_t* x = a + (uint64_t)blk * len;
                for (uint32_t j = 0; j < half; ++j) {
                    uint32_t u = x[j], v = x[j + half];
                    x[j] = p2_add(u, v);
                    x[j + half] = p2_mul_shoup(p2_sub(u, v), tw[(uint64_t)j * stride], twq[(uint64_t)j * stride]);
                }
            }
        }
    } else {
        for (uint32_t len = 2; len <= n; len <<= 1) {
            uint32_t half = len / 2, stride = n / len;
#pragma omp parallel for schedule(static)
            for (int64_t blk = 0; blk < (int64_t)(n / len); ++blk) {
                uint32_t* x = a + (uint64_t)blk * len;
                for (uint32_t j = 0; j < half; ++j) {
                    uint32_t u = x[j], v = p2_mul_shoup(x[j + half], tw[(uint64_t)j * stride], twq[(uint64_t)j * stride]);
                    x[j] = p2_add(u, v);
                    x[j + half] = p2_sub(u, v);
                }
            }
        }
    }
}
/* a <- a (*) b cyclically over Lc <= 2^30, plain residues (0/1 inputs) in and out. */
static int ntt2_cyclic(uint32_t* a, uint32_t* b, uint32_t Lc) {
    uint32_t half = Lc / 2;
    uint32_t* tw = (uint32_t*)malloc((size_t)half * 4 + 4);
    uint32_t* twq = (uint32_t*)malloc((size_t)half * 4 + 4);
    if (!tw || !twq) { free(tw); free(twq); return -1; }
    uint32_t w = p2_pow(5, (uint64_t)(NTT2_P - 1) / Lc);
    p2_tables(w, half, tw, twq);
    p2_core(a, Lc, tw, twq, 0);
    p2_core(b, Lc, tw, twq, 0);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)Lc; ++j) a[j] = p2_mul(a[j], b[j]);
    p2_tables(p2_pow(w, (uint64_t)Lc - 1), half, tw, twq);
    p2_core(a, Lc, tw, twq, 1);
    uint32_t ninv = p2_pow(Lc, (uint64_t)NTT2_P - 2);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)Lc; ++j) a[j] = p2_mul(a[j], ninv);
    free(tw); free(twq);
    return 0;
}

/* counts[p] = sum_{j < kbits} x[j] s[(p - j) mod n] exactly, p < nout.  Power-of-two n: one cyclic
 * NTT of length n.  Other n: the linear convolution of x with the seed repeated twice (length 2n)
 * read at p + n, on an NTT of length >= kbits + 2n so nothing wraps.  Exact while every sum stays
 * below the prime (kbits < 2^31); lengths up to 2^27 use 15*2^27+1, longer ones 3*2^30+1. */
int pa_or_conv_counts(const uint32_t* ks, uint32_t kbits, const uint32_t* seed, uint32_t n,
                      uint32_t nout, uint32_t* counts) {
    if (n < 2 || kbits > n || nout > n) return -1;
    const int pow2 = (n & (n - 1)) == 0;
    uint64_t Lc = 1;
    while (Lc < (pow2 ? (uint64_t)n : (uint64_t)kbits + 2ull * n)) Lc <<= 1;
    if (Lc > (1ull << 30)) return -1;
    const int big = Lc > (1ull << 27) || ntt_force_big;  /* beyond 15*2^27+1's roots: 3*2^30+1 */
    ntt_init();
    uint32_t* a = (uint32_t*)calloc((size_t)Lc, 4);
    uint32_t* b = (uint32_t*)calloc((size_t)Lc, 4);
    if (!a || !b) { free(a); free(b); return -1; }
    uint32_t one = big ? 1u : to_mont(1);
    const uint64_t bl = pow2 ? n : 2ull * n;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)kbits; ++j) a[j] = refbit(ks, (uint64_t)j) ? one : 0u;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)bl; ++t) b[t] = refbit(seed, (uint64_t)t % n) ? one : 0u;
    int rc = big ? ntt2_cyclic(a, b, (uint32_t)Lc) : ntt_cyclic(a, b, (uint32_t)Lc);
    if (rc == 0) {
        const uint64_t base = pow2 ? 0 : n;
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < (int64_t)nout; ++p) counts[p] = big ? a[base + (uint64_t)p] : from_mont(a[base + (uint64_t)p]);
    }
    free(a); free(b);
    return rc;
}

static int conv_ntt(const uint32_t* ks, uint32_t kbits, const uint32_t* seed, uint32_t n,
                    uint32_t nout, uint8_t* out) {
    uint32_t* c = (uint32_t*)malloc((size_t)(nout ? nout : 1) * 4);
    if (!c) return -1;
    int rc = pa_or_conv_counts(ks, kbits, seed, n, nout, c);
    if (rc == 0) {
        memset(out, 0, (nout + 7) / 8);
#pragma omp parallel for schedule(static)
        for (int64_t byte = 0; byte < (int64_t)((nout + 7) / 8); ++byte) {
            uint8_t v = 0;
            for (uint32_t i = 0; i < 8; ++i) {
                uint64_t p = (uint64_t)byte * 8 + i;
                if (p < nout && (c[p] & 1u)) v |= (uint8_t)(0x80u >> i);
            }
            out[byte] = v;
        }
    }
    free(c);
    return rc;
}

/* ---------------------------------------------------------------- FFT restatement ---- */
/* Radix-2 complex FFT (interleaved re/im), FP64 twiddles rounded to the working type.  The
 * first log2(BLK) stages run block by block so they stay in cache. */
#define FFT_BLK_LOG 14
#define DEFINE_FFT(T, SUF)                                                                        \
    static void bitrev_copy_##SUF(const T* in, T* out, uint32_t n) {                              \
        int lg = 0; while ((1u << lg) < n) ++lg;                                                  \
        _Pragma("omp parallel for schedule(static)")                                              \
        for (int64_t i = 0; i < (int64_t)n; ++i) {                                                \
            uint32_t r = lg ? (bitrev32((uint32_t)i) >> (32 - lg)) : 0;             \
            out[2 * (uint64_t)r] = in[2 * i]; out[2 * (uint64_t)r + 1] = in[2 * i + 1];            \
        }                                                                                         \
    }                                                                                             \
    /* sign=-1 forward (exp(-2 pi i/n)), +1 inverse; unnormalized; in: natural order */          \
    static void fft_##SUF(T* data, T* tmp, uint32_t n, int sign) {                                \
        bitrev_copy_##SUF(data, tmp, n);                                                          \
        T* tw = (T*)malloc(sizeof(T) * (n > 1 ? n : 2));                                           \
        _Pragma("omp parallel for schedule(static)")                                              \
        for (int64_t j = 0; j < (int64_t)(n / 2); ++j) {                                          \
            double ang = -2.0 * M_PI * (double)j / (double)n;                                     \
            tw[2 * j] = (T)cos(ang); tw[2 * j + 1] = (T)(sign < 0 ? sin(ang) : -sin(ang));        \
        }                                                                                         \
        uint32_t blk = n < (1u << FFT_BLK_LOG) ? n : (1u << FFT_BLK_LOG);                         \
        _Pragma("omp parallel for schedule(static)")                                              \
        for (int64_t b0 = 0; b0 < (int64_t)n; b0 += blk) {                                        \
            T* x = tmp + 2 * b0;                                                                  \
            for (uint32_t len = 2; len <= blk; len <<= 1) {                                       \
                uint32_t half = len / 2, stride = n / len;                                        \
                for (uint32_t s = 0; s < blk; s += len)                                           \
                    for (uint32_t j = 0; j < half; ++j) {                                         \
                        T wr = tw[2 * (uint64_t)j * stride], wi = tw[2 * (uint64_t)j * stride + 1]; \
                        T* u = x + 2 * (s + j); T* v = x + 2 * (s + j + half);                    \
                        T vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;                 \
                        v[0] = u[0] - vr; v[1] = u[1] - vi; u[0] += vr; u[1] += vi;               \
                    }                                                                             \
            }                                                                                     \
        }                                                                                         \
        for (uint32_t len = blk * 2; len <= n; len <<= 1) {                                       \
            uint32_t half = len / 2, stride = n / len;                                            \
            _Pragma("omp parallel for schedule(static)")                                          \
            for (int64_t q = 0; q < (int64_t)(n / 2); ++q) {                                      \
                uint64_t s = ((uint64_t)q / half) * len, j = (uint64_t)q % half;                  \
                T wr = tw[2 * j * stride], wi = tw[2 * j * stride + 1];                           \
                T* u = tmp + 2 * (s + j); T* v = tmp + 2 * (s + j + half);                        \
                T vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;                         \
                v[0] = u[0] - vr; v[1] = u[1] - vi; u[0] += vr; u[1] += vi;                       \
            }                                                                                     \
        }                                                                                         \
        memcpy(data, tmp, sizeof(T) * 2 * (size_t)n);                                             \
        free(tw);                                                                                 \
    }                                                                                             \
    /* R2C of n reals (n even) via the half-length complex transform; out: n/2+1 complex */       \
    static void r2c_##SUF(const T* x, T* out, T* tmp, uint32_t n) {                               \
        uint32_t N = n / 2;                                                                       \
        T* z = out; /* reuse out[0..N) as the packed sequence */                                  \
        memcpy(z, x, sizeof(T) * n);                                                              \
        fft_##SUF(z, tmp, N, -1);                                                                 \
        memcpy(tmp, z, sizeof(T) * 2 * (size_t)N);                                                \
        _Pragma("omp parallel for schedule(static)")                                              \
        for (int64_t k = 0; k <= (int64_t)N; ++k) {                                               \
            uint64_t k0 = (uint64_t)k % N, k1 = (N - (uint64_t)k % N) % N;                        \
            T ar = tmp[2 * k0], ai = tmp[2 * k0 + 1], br = tmp[2 * k1], bi = -tmp[2 * k1 + 1];    \
            T er = (T)0.5 * (ar + br), ei = (T)0.5 * (ai + bi);                                   \
            T dr = (T)0.5 * (ar - br), di = (T)0.5 * (ai - bi);                                   \
            double ang = -2.0 * M_PI * (double)k / (double)n;                                     \
            T wr = (T)cos(ang), wi = (T)sin(ang);                                                 \
            /* X = E - i*w*D */                                                                   \
            T tr = wr * dr - wi * di, ti = wr * di + wi * dr;                                     \
            out[2 * k] = er + ti; out[2 * k + 1] = ei - tr;                                       \
        }                                                                                         \
    }                                                                                             \
    /* C2R: in n/2+1 complex (Hermitian half), out n reals, unnormalized like cufftExecC2R */    \
    static void c2r_##SUF(const T* in, T* out, T* tmp, uint32_t n) {                              \
        uint32_t N = n / 2;                                                                       \
        _Pragma("omp parallel for schedule(static)")                                              \
        for (int64_t k = 0; k < (int64_t)N; ++k) {                                                \
            T ar = in[2 * k], ai = in[2 * k + 1], br = in[2 * (N - k)], bi = -in[2 * (N - k) + 1]; \
            T er = (T)0.5 * (ar + br), ei = (T)0.5 * (ai + bi);                                   \
            T dr = (T)0.5 * (ar - br), di = (T)0.5 * (ai - bi);                                   \
            double ang = 2.0 * M_PI * (double)k / (double)n;                                      \
            T wr = (T)cos(ang), wi = (T)sin(ang);                                                 \
            /* Q = E + i*conj(w)*D ; conj(w) = exp(+2 pi i k/n) */                                \
            T tr = wr * dr - wi * di, ti = wr * di + wi * dr;                                     \
            out[2 * k] = er - ti; out[2 * k + 1] = ei + tr;                                       \
        }                                                                                         \
        fft_##SUF(out, tmp, N, +1);                                                               \
        _Pragma("omp parallel for schedule(static)")                                              \
        for (int64_t j = 0; j < (int64_t)n; ++j) out[j] *= (T)2;                                  \
    }
DEFINE_FFT(float, f32)
DEFINE_FFT(double, f64)

/* The seed half of the reference's FP32 pipeline (binInt2float and cufftExecR2C of the seed,
 * PrivacyAmplification.cu:1819-1830, 1903; DC zeroed as setFirstElementToZero does at :1946):
 * S = n/2+1 complex (n+2 floats), *ws = the seed's popcount.  The reference computes it once
 * and reuses it while a seed is reused (reuse_seed_amount != 0, :1781-1791, 1898-1910). */
int pa_or_fft32_seed_spectrum(const uint32_t* seed, uint32_t n, float* S, uint32_t* ws) {
    const uint32_t R = 11;
    float* seed_f = (float*)calloc((size_t)n + 2, 4);
    float* tmp = (float*)malloc(sizeof(float) * ((size_t)n + 2));
    if (!seed_f || !tmp) { free(seed_f); free(tmp); return -1; }
    *ws = pa_or_bin_int2float(seed, n, R, seed_f);
    r2c_f32(seed_f, S, tmp, n);
    S[0] = S[1] = 0.0f;
    free(seed_f); free(tmp);
    return 0;
}

/* The reference's FP32 pipeline up to the float values ToBinaryArray consumes.
 * Returns v (n floats, cuFFT-style unnormalized C2R output), corr and norm.  S_static != NULL:
 * the seed's spectrum from pa_or_fft32_seed_spectrum (static seed), ws_static its popcount. */
static int fft32_pipeline_s(const uint32_t* ks, uint32_t kbits, const uint32_t* seed, uint32_t n,
                            const float* S_static, uint32_t ws_static, float* v, float* corr_out,
                            float* norm_out) {
    const uint32_t R = 11, Q = 5; /* reduction_exp, pre_mul_reduction_exp (config.yaml:42,47) */
    float* key_f = (float*)calloc((size_t)n + 2, 4);
    float* X = (float*)malloc(sizeof(float) * ((size_t)n + 2));
    float* S = S_static ? NULL : (float*)malloc(sizeof(float) * ((size_t)n + 2));
    float* tmp = (float*)malloc(sizeof(float) * ((size_t)n + 2));
    if (!key_f || !X || (!S_static && !S) || !tmp) {
        free(key_f); free(X); free(S); free(tmp); return -1;
    }
    uint32_t ws = ws_static;
    if (!S_static && pa_or_fft32_seed_spectrum(seed, n, S, &ws) != 0) {
        free(key_f); free(X); free(S); free(tmp); return -1;
    }
    uint32_t wk = pa_or_bin_int2float(ks, kbits, R, key_f);
    *corr_out = pa_or_correction(wk, ws, n);
    r2c_f32(key_f, X, tmp, n);
    X[0] = X[1] = 0.0f;                                           /* setFirstElementToZero */
    pa_or_element_wise_product(X, S_static ? S_static : S, n / 2 + 1, (float)(1u << Q)); /* ElementWiseProduct */
    c2r_f32(X, v, tmp, n);
    const float total_reduction = (float)((1u << R) * (1u << Q));
    *norm_out = ((float)n) / total_reduction / total_reduction;    /* PrivacyAmplification.cu:1729-1730 */
    free(key_f); free(X); free(S); free(tmp);
    return 0;
}
static int fft32_pipeline(const uint32_t* ks, uint32_t kbits, const uint32_t* seed, uint32_t n,
                          float* v, float* corr_out, float* norm_out) {
    return fft32_pipeline_s(ks, kbits, seed, n, NULL, 0, v, corr_out, norm_out);
}

/* The same convolution in FP64 with 0/1 inputs; v[p] = conv + corr with the DC removed, which is
 * what an exact evaluation of the reference's rounding argument gives. */
static int fft64_values(const uint32_t* ks, uint32_t kbits, const uint32_t* seed, uint32_t n,
                        double* v /* n */) {
    double* xf = (double*)calloc((size_t)n + 2, 8);
    double* sf = (double*)calloc((size_t)n + 2, 8);
    double* X = (double*)malloc(sizeof(double) * ((size_t)n + 2));
    double* S = (double*)malloc(sizeof(double) * ((size_t)n + 2));
    double* tmp = (double*)malloc(sizeof(double) * ((size_t)n + 2));
    if (!xf || !sf || !X || !S || !tmp) { free(xf); free(sf); free(X); free(S); free(tmp); return -1; }
    uint64_t wk = 0, ws = 0;
    for (uint32_t j = 0; j < kbits; ++j) { xf[j] = refbit(ks, j); wk += (uint64_t)xf[j]; }
    for (uint32_t j = 0; j < n; ++j) { sf[j] = refbit(seed, j); ws += (uint64_t)sf[j]; }
    r2c_f64(xf, X, tmp, n);
    r2c_f64(sf, S, tmp, n);
    X[0] = X[1] = 0.0; S[0] = S[1] = 0.0;
    for (uint32_t k = 0; k <= n / 2; ++k) {
        double ar = X[2 * k], ai = X[2 * k + 1], br = S[2 * k], bi = S[2 * k + 1];
        X[2 * k] = ar * br - ai * bi; X[2 * k + 1] = ar * bi + ai * br;
    }
    c2r_f64(X, v, tmp, n);
    double corr = fmod((double)(wk * ws % (2ull * n)) / (double)n, 2.0);
    for (uint32_t j = 0; j < n; ++j) v[j] = v[j] / (double)n + corr;
    free(xf); free(sf); free(X); free(S); free(tmp);
    return 0;
}

int pa_or_fft_values(const uint32_t* ks, uint32_t kbits, const uint32_t* seed, uint32_t n,
                     uint32_t nout, int use_double, double* values) {
    if (n < 4 || (n & (n - 1)) || kbits > n || nout > n) return -1;
    if (use_double) {
        double* v = (double*)malloc(sizeof(double) * (size_t)n);
        if (!v) return -1;
        int rc = fft64_values(ks, kbits, seed, n, v);
        if (rc == 0) memcpy(values, v, sizeof(double) * nout);
        free(v);
        return rc;
    }
    float* v = (float*)malloc(sizeof(float) * ((size_t)n + 2));
    if (!v) return -1;
    float corr, norm;
    int rc = fft32_pipeline(ks, kbits, seed, n, v, &corr, &norm);
    if (rc == 0)
        for (uint32_t p = 0; p < nout; ++p) values[p] = (double)(v[p] / norm + corr);
    free(v);
    return rc;
}

static int conv_fft(const uint32_t* ks, uint32_t kbits, const uint32_t* seed, uint32_t n,
                    uint32_t nout, uint8_t* out, int use_double) {
    double* v = (double*)malloc(sizeof(double) * ((size_t)nout + 1));
    if (!v) return -1;
    int rc = pa_or_fft_values(ks, kbits, seed, n, nout, use_double, v);
    if (rc == 0) {
        memset(out, 0, (nout + 7) / 8);
        for (uint32_t p = 0; p < nout; ++p) {
            long r = use_double ? lrint(v[p]) : lrintf((float)v[p]);
            if (r & 1) out[p >> 3] |= (uint8_t)(0x80u >> (p & 7));
        }
    }
    free(v);
    return rc;
}

int pa_or_conv_bits(const uint32_t* ks, uint32_t kbits, const uint32_t* seed, uint32_t n,
                    uint32_t nout, uint8_t* out, int method) {
    if (n < 32 || (n % 32) || kbits > n || nout > n) return -1;
    switch (method) {
    case PA_OR_DIRECT: return conv_direct(ks, kbits, seed, n, nout, out);
    case PA_OR_NTT: return conv_ntt(ks, kbits, seed, n, nout, out);
    case PA_OR_FFT32: return conv_fft(ks, kbits, seed, n, nout, out, 0);
    case PA_OR_FFT64: return conv_fft(ks, kbits, seed, n, nout, out, 1);
    default: return -1;
    }
}

/* ---------------------------------------------------------------- one block ---- */
static int64_t amp_out_impl(const uint32_t* key, const uint32_t* seed, uint32_t n, uint32_t m,
                            int do_xor, int do_compress, int reverse_endian, uint8_t* out, int method,
                            const float* S_static, uint32_t ws_static) {
    if (n < 64 || (n % 32) || m == 0 || (m % 32) || m >= n) return -1;
    uint32_t h = n - m, kbits = h + 1;
    uint32_t out_bits = do_compress ? m : n / 2;
    /* ToBinaryArray XORs the key rest into every word it writes (PrivacyAmplification.cu:753-777,
     * launched over vertical_block words at :2029); sendData sends out_bits of them (:1186-1197).
     * Without compression and m > n/2 only the first n/2 output bits exist, so the XOR stops
     * there; with m < n/2 the key rest ends at bit m. */
    uint32_t xor_bits = m < out_bits ? m : out_bits;
    uint32_t nw = n / 32;
    uint32_t* ks = (uint32_t*)calloc(nw + 1, 4);
    uint32_t* kr = (uint32_t*)calloc(nw + 1, 4);
    uint8_t* conv = (uint8_t*)calloc((out_bits + 7) / 8 + 4, 1);
    if (!ks || !kr || !conv) { free(ks); free(kr); free(conv); return -1; }
    if (do_xor) {
        pa_or_key_split(key, n, m, ks, kr);
    } else {
        /* reciveDataKey NoXOR branch (PrivacyAmplification.cu:1068-1080): h/32+1 words, last masked */
        memcpy(ks, key, (size_t)(h / 32) * 4);
        ks[h / 32] = key[h / 32] & 0x80000000u;
    }
    int rc;
    if (method == PA_OR_FFT32) {
        /* ToBinaryArray consumes the float values directly. */
        float* v = (float*)malloc(sizeof(float) * ((size_t)n + 2));
        float corr, norm;
        rc = v ? fft32_pipeline_s(ks, kbits, seed, n, S_static, ws_static, v, &corr, &norm) : -1;
        if (rc == 0) {
            uint32_t nwo = (out_bits + 31) / 32;
            uint32_t* words = (uint32_t*)calloc(nwo, 4);
            uint32_t* krx = (uint32_t*)calloc(nwo, 4);
            if (!words || !krx) rc = -1;
            else {
                if (do_xor) memcpy(krx, kr, (size_t)(xor_bits / 32) * 4);
                pa_or_to_binary_array(v, out_bits, norm, corr, do_xor ? krx : NULL, reverse_endian, words);
                memcpy(out, words, reverse_endian ? (out_bits + 7) / 8 : (size_t)nwo * 4);
            }
            free(words); free(krx);
        }
        free(v);
    } else {
        rc = pa_or_conv_bits(ks, kbits, seed, n, out_bits, conv, method);
        if (rc == 0) {
            if (do_xor)
                for (uint32_t p = 0; p < xor_bits; ++p)
                    if (refbit(kr, p)) conv[p >> 3] ^= (uint8_t)(0x80u >> (p & 7));
            if (reverse_endian) {
                memcpy(out, conv, (out_bits + 7) / 8);
            } else {
                uint32_t nwo = (out_bits + 31) / 32;
                for (uint32_t w = 0; w < nwo; ++w) {
                    uint32_t word = 0;
                    for (int b = 0; b < 4; ++b) word = (word << 8) | conv[4 * w + b];
                    memcpy(out + 4 * w, &word, 4);
                }
            }
        }
    }
    free(ks); free(kr); free(conv);
    return rc == 0 ? (int64_t)out_bits : -1;
}

int64_t pa_or_amp_out(const uint32_t* key, const uint32_t* seed, uint32_t n, uint32_t m,
                      int do_xor, int do_compress, int reverse_endian, uint8_t* out, int method) {
    return amp_out_impl(key, seed, n, m, do_xor, do_compress, reverse_endian, out, method, NULL, 0);
}

int64_t pa_or_amp_out_fft32_seeded(const uint32_t* key, const float* S, uint32_t ws, uint32_t n, uint32_t m,
                                   int do_xor, int do_compress, int reverse_endian, uint8_t* out) {
    if (!S) return -1;
    return amp_out_impl(key, NULL, n, m, do_xor, do_compress, reverse_endian, out, PA_OR_FFT32, S, ws);
}

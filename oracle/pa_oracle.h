/*
 * pa_oracle.h -- CPU oracle for the Toeplitz-hashing path of nicoboss/PrivacyAmplification.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (privacyamplification_amd/) links, loads or
 * calls this library.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * use it, and only as the checker (or, for the baseline, as the thing timed on the host CPU).
 *
 * Bit layouts (same as the reference's wire format):
 *   RefBits  : bit p of a stream is (w[p>>5] >> (31-(p&31))) & 1, w = little-endian uint32 words
 *              (PrivacyAmplification.cu:223-257, 584-593; LargeBlocksizeExample/main.cpp:220-230).
 *   OutBytes : output bit p is byte p>>3, bit 7-(p&7) (plain MSB-first byte string); this is the
 *              AMPOUT_REVERSE_ENDIAN=TRUE layout (PrivacyAmplification.h:38, .cu:260-267,757-762).
 *              With reverse_endian=0 the output is RefBits words instead.
 */
#ifndef PA_ORACLE_H
#define PA_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { PA_OR_DIRECT = 0, PA_OR_NTT = 1, PA_OR_FFT32 = 2, PA_OR_FFT64 = 3 };

/* xoshiro256** seeded by splitmix64; Bernoulli(p) bits packed as RefBits.  For p == 0.5 each
 * 64-bit output supplies two words (high half first); otherwise one output per bit with
 * bit = ((x >> 11) * 2^-53) < p.  Words beyond nbits are zero-filled. */
void pa_or_synth_bits(uint64_t seed, double p, uint64_t nbits, uint32_t* out_words);

/* key2StartRest, PrivacyAmplification.cu:861-892.  key: n/32+1 words (n+1 bits).
 * key_start: n/32 words (bits 0..h, word h/32 masked to its MSB, rest zero).
 * key_rest : m/32 words (bits h+1..h+m, a 1-bit left shift across words). */
void pa_or_key_split(const uint32_t* key, uint32_t n, uint32_t m, uint32_t* key_start,
                     uint32_t* key_rest);

/* binInt2float, PrivacyAmplification.cu:570-602: RefBits -> {0, 2^-R} floats; returns popcount. */
uint32_t pa_or_bin_int2float(const uint32_t* bits, uint32_t nbits, uint32_t reduction_exp,
                             float* out);

/* calculateCorrectionFloat, PrivacyAmplification.cu:363-373: (float)fmod((u32)(wk*ws)/(double)n, 2). */
float pa_or_correction(uint32_t wk, uint32_t ws, uint32_t n);

/* ElementWiseProduct, PrivacyAmplification.cu:465-481: a[i] = (a[i]/r)*(b[i]/r) (complex). */
void pa_or_element_wise_product(float* a, const float* b, uint32_t count, float r);

/* ToBinaryArray / ToBinaryArrayNoXOR, PrivacyAmplification.cu:739-813: bits of
 * rint(v/norm + corr) & 1, packed with the OutBytes (or RefBits) shift table, XORed with the
 * byte-swapped key_rest words when key_rest != NULL.  Writes ceil(nbits/32) words. */
void pa_or_to_binary_array(const float* v, uint32_t nbits, float norm, float corr,
                           const uint32_t* key_rest, int reverse_endian, uint32_t* out_words);

/* Circular GF(2) product: conv[p] = XOR_{j<kbits} x[j] & s[(p-j) mod n] for p < nout,
 * x = first kbits bits of key_start (RefBits), s = n-bit seed (RefBits).
 * Written as an MSB-first bit string (OutBytes layout) of ceil(nout/8) bytes.
 * method: PA_OR_DIRECT (bit-packed schoolbook), PA_OR_NTT (exact integer NTT mod 15*2^27+1, or
 * mod 3*2^30+1 for transform lengths above 2^27; any n with n <= 2^29 when n is a power of two),
 * PA_OR_FFT64 (double FFT, rounded).  Returns 0 on success. */
int pa_or_conv_bits(const uint32_t* key_start, uint32_t kbits, const uint32_t* seed, uint32_t n,
                    uint32_t nout, uint8_t* out_bytes, int method);

/* One block with the reference's wire semantics (mainloop body, PrivacyAmplification.cu:1819-2042):
 *   do_xor=1 : key has n/32+1 words, split by key2StartRest; out = m bits of T*x ^ key_rest.
 *   do_xor=0 : key has h/32+1 words (already key_start); out = m bits (compress) or n/2 bits.
 *   do_compress=0 with do_xor=1 outputs n/2 bits, the first m XORed with key_rest.
 * method PA_OR_FFT32 restates the reference's FP32 float pipeline (binInt2float, R2C, DC zero,
 * ElementWiseProduct, C2R, ToBinaryArray) with reduction_exp=11, pre_mul_reduction_exp=5.
 * out: ceil(out_bits/8) bytes (OutBytes) or ceil(out_bits/32) words when reverse_endian=0.
 * Returns the number of output bits, or -1 on bad arguments. */
int64_t pa_or_amp_out(const uint32_t* key, const uint32_t* seed, uint32_t n, uint32_t m,
                      int do_xor, int do_compress, int reverse_endian, uint8_t* out, int method);

/* Static-seed form of the FP32 restatement: the seed's half spectrum S (n+2 floats, DC zeroed)
 * and popcount ws are computed once by pa_or_fft32_seed_spectrum, and each block then runs only
 * the key's transform, the product and the inverse (reuse_seed_amount != 0,
 * PrivacyAmplification.cu:1781-1791, 1898-1910).  Same output as pa_or_amp_out(..., PA_OR_FFT32). */
int pa_or_fft32_seed_spectrum(const uint32_t* seed, uint32_t n, float* S, uint32_t* ws);
int64_t pa_or_amp_out_fft32_seeded(const uint32_t* key, const float* S, uint32_t ws, uint32_t n, uint32_t m,
                                   int do_xor, int do_compress, int reverse_endian, uint8_t* out);

/* Pre-rounding values of the FP32 (or FP64) FFT restatement: v[p] = conv/norm + corr for p < nout,
 * i.e. the numbers ToBinaryArray rounds.  Used to measure precision margins. */
int pa_or_fft_values(const uint32_t* key_start, uint32_t kbits, const uint32_t* seed, uint32_t n,
                     uint32_t nout, int use_double, double* values);

/* Exact integer circular convolution c[p] = sum_{j<kbits} x[j] s[(p-j) mod n] for p < nout
 * (n a power of two <= 2^30, or any n with kbits + 2n <= 2^30). */
int pa_or_conv_counts(const uint32_t* key_start, uint32_t kbits, const uint32_t* seed, uint32_t n,
                      uint32_t nout, uint32_t* counts);

/* Test hook: 1 makes every NTT use the prime 3*2^30+1 (the long-length path), 0 restores. */
void pa_or_set_ntt_big(int on);

/* Threads used by the OpenMP regions (0 = library default). */
void pa_or_set_threads(int nthreads);
int pa_or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif

// pa_kernels.hip -- gfx950 kernels of the Toeplitz-hashing engine and their launchers.
//
// The reference evaluates amp_out with cuFFT/VkFFT and five small kernels, one synchronised
// launch each (binInt2float, calculateCorrectionFloat, setFirstElementToZero, ElementWiseProduct,
// ToBinaryArray; PrivacyAmplification.cu:363-813, mainloop 1819-2042).  Here the same arithmetic
// (FP32 FFT convolution, DC-zeroed spectra, round-mod-2 with the popcount correction) is fused
// into a few HBM passes:
//
//   k_transpose_bits  key/seed RefBits -> per-(class,column) bit masks for pass A, plus popcounts
//                     (binInt2float's popcount, PrivacyAmplification.cu:570-602)
//   k_fwd_cols        pass A: decimated column FFTs of the +-1/2-encoded bits (R2C forward part)
//   k_rows            pass B: row FFTs, real split, DC zeroing (setFirstElementToZero, :408-418),
//                     complex product (ElementWiseProduct, :465-481), inverse split, inverse row FFTs
//   k_inv_cols        pass A': inverse column FFTs + round-mod-2 with the correction term
//                     (calculateCorrectionFloat :363-373, ToBinaryArray :739-783)
//   k_pack            output word assembly with the key-rest XOR and byte order
//                     (key2StartRest's key_rest :873-878, ToBinaryArray's bswap XOR :757-762)
//   k_small           the whole block in one workgroup for L <= 2^14
#include <hip/hip_runtime.h>

#include <cstdio>

#include "pa_device.hpp"
#include "pa_geom.hpp"
#include "pa_kernels.hpp"

namespace pa {

// ------------------------------------------------------------------ logical bit sources ----
// Word widx of the key sequence x: bits 0..h of the key (key2StartRest's key_start,
// PrivacyAmplification.cu:867-868: word h/32 keeps only its MSB), zero beyond.
__device__ __forceinline__ uint32_t key_word(const uint32_t* __restrict__ key, const Geom& g, uint32_t widx) {
    if (widx < g.key_hw) return key[widx];
    if (widx == g.key_hw) return key[widx] & 0x80000000u;
    return 0u;
}
// Word widx of the seed sequence t: the n seed bits rotated left by seed_rot words, zero beyond n.
__device__ __forceinline__ uint32_t seed_word(const uint32_t* __restrict__ seed, const Geom& g, uint32_t widx) {
    if (widx >= g.seed_words) return 0u;
    uint32_t s = widx + g.seed_rot;
    if (s >= g.seed_words) s -= g.seed_words;
    return seed[s];
}

// corr = (wk * ws / L) mod 2, exactly: the mean term that zeroing both DC bins removes
// (calculateCorrectionFloat, PrivacyAmplification.cu:363-373, evaluated in 64-bit integers).
__device__ __forceinline__ float correction(uint32_t wk, uint32_t ws, uint32_t L) {
    const uint64_t prod = (uint64_t)wk * (uint64_t)ws;
    const uint64_t r = prod & (2ull * L - 1ull);  // L is a power of two
    return (float)((double)r / (double)L);
}

__device__ __forceinline__ uint32_t round_bit(float v, float& dist) {
    const float r = rintf(v);
    dist = fmaxf(dist, fabsf(v - r));
    return (uint32_t)__float2int_rn(v) & 1u;
}

__global__ __launch_bounds__(64) void k_zero_u32(uint32_t* p, uint32_t count) {
    const uint32_t i = threadIdx.x;
    if (i < count) p[i] = 0u;
}

// ------------------------------------------------------------------ bit transpose ----
// tb[u*N2 + t2] = E | O << 16 with E bit v = x[2(t2 + N2(u + M v))], O bit v = the odd bit.
template <int D>
__global__ __launch_bounds__(256) void k_transpose_bits(const uint32_t* __restrict__ bits,
                                                        uint32_t* __restrict__ tb,
                                                        uint32_t* __restrict__ count, Geom g,
                                                        int is_seed) {
    const uint32_t total = g.M * g.N2;
    const uint32_t idx = blockIdx.x * 256u + threadIdx.x;
    uint32_t pc = 0;
    if (idx < total) {
        const uint32_t t2 = idx & (g.N2 - 1u), u = idx >> g.lg2;
        uint32_t E = 0, O = 0;
#pragma unroll
        for (int v = 0; v < D; ++v) {
            const uint32_t pos = 2u * (t2 + g.N2 * (u + g.M * (uint32_t)v));
            const uint32_t widx = pos >> 5, sh = pos & 31u;
            const uint32_t w = is_seed ? seed_word(bits, g, widx) : key_word(bits, g, widx);
            E |= ((w >> (31u - sh)) & 1u) << v;
            O |= ((w >> (30u - sh)) & 1u) << v;
        }
        tb[idx] = E | (O << 16);
        pc = (uint32_t)(__popc(E) + __popc(O));
    }
    for (int off = 32; off > 0; off >>= 1) pc += __shfl_xor(pc, off);
    if ((threadIdx.x & 63u) == 0u && pc) atomicAdd(count, pc);
}

// ------------------------------------------------------------------ pass A ----
// Workgroup (column group grp of 16 columns, class c): computes Y[k1][t2] for k1 = c + D*k1',
// t2 in the group:  Y = w_N^{t2 k1} * FFT_M( w_N1^{u c} * sum_v z[t2 + N2(u + M v)] w_D^{v c} ).
template <int LGM, int D>
__global__ __launch_bounds__(256) void k_fwd_cols(const uint32_t* __restrict__ tb,
                                                  float2* __restrict__ Y, Geom g,
                                                  const float2* __restrict__ tw14,
                                                  const float2* __restrict__ twlo,
                                                  const float2* __restrict__ twhi) {
    constexpr int M = 1 << LGM;
    constexpr int NT = 256;
    constexpr int SEQS = seq_stride(M);
    __shared__ float2 buf[16 * SEQS];
    __shared__ float2 TE[256];
    const uint32_t logical = xcd_logical(blockIdx.x, gridDim.x);
    const uint32_t c = logical % D, grp = logical / D;
    constexpr int VB = D < 8 ? D : 8;
    for (int b = threadIdx.x; b < 256; b += NT) {
        float2 acc = make_float2(0.f, 0.f);
        for (int v = 0; v < VB; ++v)
            if ((b >> v) & 1) acc = cadd(acc, tw14[((v * c) % D) * (TW14 / D)]);
        TE[b] = acc;
    }
    const float2 w8c = (D > 8) ? tw14[((8u * c) % D) * (TW14 / D)] : make_float2(1.f, 0.f);
    __syncthreads();
    const uint32_t step_n1 = TW14 / g.N1;
    for (int e = threadIdx.x; e < 16 * M; e += NT) {
        const int col = e & 15, u = e >> 4;
        const uint32_t t = tb[(uint32_t)u * g.N2 + grp * 16u + (uint32_t)col];
        const uint32_t E = t & 0xFFFFu, O = t >> 16;
        float2 sE = TE[E & 255u], sO = TE[O & 255u];
        if constexpr (D > 8) {
            sE = cadd(sE, cmul(w8c, TE[E >> 8]));
            sO = cadd(sO, cmul(w8c, TE[O >> 8]));
        }
        float2 z = make_float2(sE.x - sO.y, sE.y + sO.x);
        if (c == 0) {
            z.x -= 0.5f * (float)D;
            z.y -= 0.5f * (float)D;
        }
        if (c != 0) z = cmul(z, tw14[((uint32_t)u * c) * step_n1]);
        buf[col * SEQS + padi(u)] = z;
    }
    __syncthreads();
    lds_fft<LGM, 16, SEQS, NT>(buf, tw14);
    for (int e = threadIdx.x; e < 16 * M; e += NT) {
        const int col = e & 15, k1p = e >> 4;
        const uint32_t k1 = c + (uint32_t)D * (uint32_t)k1p, t2 = grp * 16u + (uint32_t)col;
        float2 y = buf[col * SEQS + padi(k1p)];
        y = cmul(y, tw_two_level(twlo, twhi, 2u * t2 * k1));
        Y[(size_t)k1 * g.N2 + t2] = y;
    }
}

// ------------------------------------------------------------------ pass B ----
// Workgroup w owns rows (w, N1-w), or (0, N1/2) for w = 0: the rows whose frequencies pair
// k <-> N-k in the real split.  After the forward row FFTs each pair (k, N-k) is one work unit:
// real split of both operands, DC zeroing, product, inverse split, written back in place, so
// nothing is held in registers across a barrier.
//   ROWS_SEEDPREP: rows of the seed's pass-A output -> split seed spectrum S (in place, plus the
//                  Nyquist value S[N] at index N).
//   ROWS_STATIC:   rows of the key's pass-A output and the resident S -> R'[k1][t2] =
//                  w_N^{t2 k1} * FFT_N2(conj Q)[t2], written in place over the key rows.
// The split twiddle w_L^k with k = row + N1*k2 factors as w_L^row * w_{2 N2}^k2 = w_L^row * tw14[k2].
template <int MODE>
__global__ __launch_bounds__(ROWS_NT) void k_rows(float2* __restrict__ Yk,
                                                  float2* __restrict__ Sst, Geom g,
                                                  const float2* __restrict__ tw14,
                                                  const float2* __restrict__ twlo,
                                                  const float2* __restrict__ twhi) {
    constexpr int LG2 = ROW_LOG, N2 = 1 << LG2, NT = ROWS_NT, SEQS = seq_stride(N2);
    static_assert(2 * N2 == TW14, "row split twiddle assumes N2 = 2^13");
    constexpr int PER = 2 * N2 / NT, HALF = PER / 2;
    __shared__ float2 buf[2 * SEQS];
    const uint32_t w = blockIdx.x, N1 = g.N1;
    const uint32_t rows[2] = {w, (w == 0) ? (N1 >> 1) : (N1 - w)};
    const int tid = threadIdx.x;
    const float2 wrow[2] = {tw_two_level(twlo, twhi, rows[0]), tw_two_level(twlo, twhi, rows[1])};
    float2* __restrict__ data = (MODE == ROWS_SEEDPREP) ? Sst : Yk;

#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int seq = i / HALF, k2 = tid + (i % HALF) * NT;
        buf[seq * SEQS + padi(k2)] = data[(size_t)rows[seq] * N2 + k2];
    }
    __syncthreads();
    lds_fft<LG2, 2, SEQS, NT>(buf, tw14);

    // One element pair: (seqA, kA2) holds frequency k, (seqB, kB2) frequency N-k.
    auto unit = [&](int seqA, int kA2, int seqB, int kB2) {
        const int oA = seqA * SEQS + padi(kA2), oB = seqB * SEQS + padi(kB2);
        const float2 zA = buf[oA], zB = buf[oB];
        const float2 wA = cmul(wrow[seqA], tw14[kA2]);
        const float2 wB = cmul(wrow[seqB], tw14[kB2]);
        const float2 xA = r2c_split(zA, zB, wA);
        const float2 xB = r2c_split(zB, zA, wB);
        const size_t gA = (size_t)rows[seqA] * N2 + kA2, gB = (size_t)rows[seqB] * N2 + kB2;
        if constexpr (MODE == ROWS_SEEDPREP) {
            Sst[gA] = xA;
            Sst[gB] = xB;
        } else {
            const float2 pA = cmul(xA, Sst[gA]);
            const float2 pB = cmul(xB, Sst[gB]);
            buf[oA] = conjf2(c2r_split(pA, pB, wA));
            buf[oB] = conjf2(c2r_split(pB, pA, wB));
        }
    };

#pragma unroll 1
    for (int p = tid; p < N2; p += NT) {
        if (w != 0) {
            unit(0, p, 1, N2 - 1 - p);
        } else if (p >= N2 / 2) {
            const int q = p - N2 / 2;
            unit(1, q, 1, N2 - 1 - q);
        } else if (p != 0) {
            unit(0, p, 0, N2 - p);
        } else {
            // k = 0: DC zeroed, Nyquist X[N] = Re Z0 - Im Z0 carried separately; and the
            // self-paired k = N/2 at (row 0, N2/2).
            const float2 z0 = buf[padi(0)];
            const float2 xN = make_float2(z0.x - z0.y, 0.f);
            const int oh = padi(N2 / 2);
            const float2 zh = buf[oh];
            const float2 wh = cmul(wrow[0], tw14[N2 / 2]);
            const float2 xh = r2c_split(zh, zh, wh);
            const size_t gh = (size_t)(N2 / 2);
            if constexpr (MODE == ROWS_SEEDPREP) {
                Sst[0] = make_float2(0.f, 0.f);
                Sst[(size_t)N1 * N2] = xN;
                Sst[gh] = xh;
            } else {
                const float2 pN = cmul(xN, Sst[(size_t)N1 * N2]);
                buf[padi(0)] = conjf2(c2r_split(make_float2(0.f, 0.f), pN, make_float2(1.f, 0.f)));
                const float2 ph = cmul(xh, Sst[gh]);
                buf[oh] = conjf2(c2r_split(ph, ph, wh));
            }
        }
    }
    if constexpr (MODE == ROWS_SEEDPREP) return;
    __syncthreads();
    lds_fft<LG2, 2, SEQS, NT>(buf, tw14);
#pragma unroll 2
    for (int i = 0; i < PER; ++i) {
        const int seq = i / HALF, t2 = tid + (i % HALF) * NT;
        const uint32_t row = rows[seq];
        const float2 F = buf[seq * SEQS + padi(t2)];
        Yk[(size_t)row * N2 + t2] = cmul(F, tw_two_level(twlo, twhi, 2u * (uint32_t)t2 * row));
    }
}

// ------------------------------------------------------------------ pass A' ----
// Workgroup grp owns columns [W*grp, W*grp + W): G[t1] = FFT_N1(R'[.][t2])[t1] = conj of the
// unnormalised inverse, so c[2t] = Re G / N and c[2t+1] = -Im G / N at t = t2 + N2*t1.
// Writes one nibble (4 conv bits, MSB first) per (t1, column pair) inside the output window.
template <int LG1>
__global__ __launch_bounds__(inv_cols_threads(LG1)) void k_inv_cols(
    const float2* __restrict__ Rp, uint8_t* __restrict__ nib, Geom g,
    const uint32_t* __restrict__ wk_ptr, const uint32_t* __restrict__ ws_ptr,
    const float2* __restrict__ tw14, float* __restrict__ margin) {
    constexpr int N1 = 1 << LG1;
    constexpr int W = inv_cols_width(LG1);
    constexpr int NT = inv_cols_threads(LG1);
    constexpr int SEQS = seq_stride(N1);
    __shared__ float2 buf[W * SEQS];
    const uint32_t grp = xcd_logical(blockIdx.x, gridDim.x);
    for (int e = threadIdx.x; e < W * N1; e += NT) {
        const int col = e % W, k1 = e / W;
        buf[col * SEQS + padi(k1)] = Rp[(size_t)k1 * g.N2 + grp * W + col];
    }
    __syncthreads();
    lds_fft<LG1, W, SEQS, NT>(buf, tw14);
    const float corr = correction(*wk_ptr, *ws_ptr, g.L);
    const float invN = g.inv_N;
    float dist = 0.f;
    constexpr int P = W / 2;
    for (int e = threadIdx.x; e < P * N1; e += NT) {
        const int p = e % P, t1 = e / P;
        const uint32_t t2 = grp * W + 2u * (uint32_t)p;
        const uint32_t i0 = 2u * (t2 + g.N2 * (uint32_t)t1);
        if (i0 + 4u <= g.out_off || i0 >= g.out_off + g.out_bits) continue;
        const float2 G0 = buf[(2 * p) * SEQS + padi(t1)];
        const float2 G1 = buf[(2 * p + 1) * SEQS + padi(t1)];
        const uint32_t b0 = round_bit(G0.x * invN + corr, dist);
        const uint32_t b1 = round_bit(-G0.y * invN + corr, dist);
        const uint32_t b2 = round_bit(G1.x * invN + corr, dist);
        const uint32_t b3 = round_bit(-G1.y * invN + corr, dist);
        nib[i0 >> 2] = (uint8_t)((b0 << 3) | (b1 << 2) | (b2 << 1) | b3);
    }
    if (margin) margin_publish(dist, margin);
}

// ------------------------------------------------------------------ output ----
// Output word w: conv bits [out_off + 32w, +32) from 8 nibbles, XOR key_rest word w
// ((key[h/32+w] << 1) | (key[h/32+w+1] >> 31), key2StartRest PrivacyAmplification.cu:873-878),
// stored byte-swapped for the OutBytes layout (ToBinaryArray, PrivacyAmplification.cu:757-762).
__global__ __launch_bounds__(256) void k_pack(const uint8_t* __restrict__ nib,
                                              const uint32_t* __restrict__ key,
                                              uint32_t* __restrict__ out, Geom g) {
    const uint32_t w = blockIdx.x * 256u + threadIdx.x;
    if (w >= (g.out_bits >> 5)) return;
    const uint32_t q = (g.out_off >> 2) + 8u * w;
    const uint2 nb = *reinterpret_cast<const uint2*>(nib + q);
    uint32_t cw = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) cw |= ((nb.x >> (8 * j)) & 0xFu) << (28 - 4 * j);
#pragma unroll
    for (int j = 0; j < 4; ++j) cw |= ((nb.y >> (8 * j)) & 0xFu) << (12 - 4 * j);
    if (g.do_xor && 32u * w < g.m) cw ^= (key[g.key_hw + w] << 1) | (key[g.key_hw + w + 1] >> 31);
    out[w] = g.reverse_endian ? __builtin_bswap32(cw) : cw;
}

// ------------------------------------------------------------------ small path ----
// The whole block in one workgroup: N = 2^LGN <= 8192 complex points per operand in LDS.
// sspec != nullptr: static seed (precomputed split spectrum, N+1 values, popcount *ws_ptr).
// SEEDPREP: only computes the seed's split spectrum into sspec_out and its popcount.
template <int LGN, bool SEEDPREP>
__global__ __launch_bounds__(small_threads(LGN)) void k_small(
    const uint32_t* __restrict__ key, const uint32_t* __restrict__ seed,
    const float2* __restrict__ sspec, const uint32_t* __restrict__ ws_ptr,
    uint32_t* __restrict__ out, Geom g, const float2* __restrict__ tw14,
    float* __restrict__ margin, float2* __restrict__ sspec_out, uint32_t* __restrict__ ws_out) {
    constexpr int N = 1 << LGN, L = 2 * N;
    constexpr int NT = small_threads(LGN);
    constexpr int SEQS = seq_stride(N);
    constexpr int PER = (N + NT - 1) / NT;
    __shared__ float2 buf[2 * SEQS];
    __shared__ uint32_t cnt[2];
    const int tid = threadIdx.x;
    const bool dyn = SEEDPREP || (sspec == nullptr);
    if (tid < 2) cnt[tid] = 0u;
    __syncthreads();
    uint32_t ck = 0, cs = 0;
    for (int t = tid; t < N; t += NT) {
        const uint32_t pos = 2u * (uint32_t)t, widx = pos >> 5, sh = pos & 31u;
        if (!SEEDPREP) {
            const uint32_t w = key_word(key, g, widx);
            const uint32_t e = (w >> (31u - sh)) & 1u, o = (w >> (30u - sh)) & 1u;
            ck += e + o;
            buf[padi(t)] = make_float2((float)e - 0.5f, (float)o - 0.5f);
        } else {
            buf[padi(t)] = make_float2(0.f, 0.f);
        }
        if (dyn) {
            const uint32_t w = seed_word(seed, g, widx);
            const uint32_t e = (w >> (31u - sh)) & 1u, o = (w >> (30u - sh)) & 1u;
            cs += e + o;
            buf[SEQS + padi(t)] = make_float2((float)e - 0.5f, (float)o - 0.5f);
        }
    }
    if (ck) atomicAdd(&cnt[0], ck);
    if (cs) atomicAdd(&cnt[1], cs);
    __syncthreads();
    if (SEEDPREP) {
        lds_fft<LGN, 2, SEQS, NT>(buf, tw14);  // sequence 0 is unused
    } else if (dyn) {
        lds_fft<LGN, 2, SEQS, NT>(buf, tw14);
    } else {
        lds_fft<LGN, 1, SEQS, NT>(buf, tw14);
    }
    constexpr int TWS = TW14 / L;
    float2 X[PER];
    float2 PN = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int k = tid + i * NT;
        if (k >= N) continue;
        const int kp = (N - k) & (N - 1);
        const float2 wk = tw14[k * TWS];
        float2 s, sn = make_float2(0.f, 0.f);
        if (dyn) {
            const float2 zk = buf[SEQS + padi(k)], zp = buf[SEQS + padi(kp)];
            s = r2c_split(zk, zp, wk);
            if (k == 0) {
                sn = make_float2(zk.x - zk.y, 0.f);
                s = make_float2(0.f, 0.f);
            }
        } else {
            s = sspec[k];
            if (k == 0) sn = sspec[N];
        }
        if (SEEDPREP) {
            sspec_out[k] = s;
            if (k == 0) sspec_out[N] = sn;
            continue;
        }
        const float2 zk = buf[padi(k)], zp = buf[padi(kp)];
        float2 x = r2c_split(zk, zp, wk);
        if (k == 0) {
            PN = cmul(make_float2(zk.x - zk.y, 0.f), sn);
            x = make_float2(0.f, 0.f);
        }
        X[i] = cmul(x, s);
    }
    if (SEEDPREP) {
        if (tid == 0) *ws_out = cnt[1];
        return;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int k = tid + i * NT;
        if (k < N) buf[padi(k)] = X[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int k = tid + i * NT;
        if (k >= N) continue;
        const int kp = (N - k) & (N - 1);
        const float2 pp = (k == 0) ? PN : buf[padi(kp)];
        X[i] = conjf2(c2r_split(X[i], pp, tw14[k * TWS]));
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int k = tid + i * NT;
        if (k < N) buf[padi(k)] = X[i];
    }
    __syncthreads();
    lds_fft<LGN, 1, SEQS, NT>(buf, tw14);
    const uint32_t ws = dyn ? cnt[1] : *ws_ptr;
    const float corr = correction(cnt[0], ws, (uint32_t)L);
    const float invN = g.inv_N;
    float dist = 0.f;
    const uint32_t nwords = g.out_bits >> 5;
    for (uint32_t wd = (uint32_t)tid; wd < nwords; wd += NT) {
        const uint32_t t0 = (g.out_off + 32u * wd) >> 1;
        uint32_t cw = 0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const float2 G = buf[padi((int)t0 + j)];
            cw |= round_bit(G.x * invN + corr, dist) << (31 - 2 * j);
            cw |= round_bit(-G.y * invN + corr, dist) << (30 - 2 * j);
        }
        if (g.do_xor && 32u * wd < g.m) cw ^= (key[g.key_hw + wd] << 1) | (key[g.key_hw + wd + 1] >> 31);
        out[wd] = g.reverse_endian ? __builtin_bswap32(cw) : cw;
    }
    if (margin) margin_publish(dist, margin);
}

// ================================================================== launchers ====
namespace {
template <int LGM, int D>
hipError_t launch_fwd_cols_t(const uint32_t* tb, float2* Y, const Geom& g, const Tables& t, hipStream_t s) {
    const uint32_t grid = (g.N2 / 16u) * (uint32_t)D;
    k_fwd_cols<LGM, D><<<grid, 256, 0, s>>>(tb, Y, g, t.tw14, t.lo, t.hi);
    return hipGetLastError();
}
template <int LG1>
hipError_t launch_inv_cols_t(const float2* R, uint8_t* nib, const Geom& g, const uint32_t* wk,
                             const uint32_t* ws, const Tables& t, float* margin, hipStream_t s) {
    const uint32_t grid = g.N2 / (uint32_t)inv_cols_width(LG1);
    k_inv_cols<LG1><<<grid, inv_cols_threads(LG1), 0, s>>>(R, nib, g, wk, ws, t.tw14, margin);
    return hipGetLastError();
}
template <int D>
hipError_t launch_transpose_t(const uint32_t* bits, uint32_t* tb, uint32_t* count, const Geom& g,
                              int is_seed, hipStream_t s) {
    const uint32_t total = g.M * g.N2;
    k_transpose_bits<D><<<(total + 255u) / 256u, 256, 0, s>>>(bits, tb, count, g, is_seed);
    return hipGetLastError();
}
template <int LGN>
hipError_t launch_small_t(const uint32_t* key, const uint32_t* seed, const float2* sspec,
                          const uint32_t* ws_ptr, uint32_t* out, const Geom& g, const Tables& t,
                          float* margin, float2* sspec_out, uint32_t* ws_out, hipStream_t s) {
    if (sspec_out)
        k_small<LGN, true><<<1, small_threads(LGN), 0, s>>>(key, seed, sspec, ws_ptr, out, g, t.tw14,
                                                           margin, sspec_out, ws_out);
    else
        k_small<LGN, false><<<1, small_threads(LGN), 0, s>>>(key, seed, sspec, ws_ptr, out, g, t.tw14,
                                                            margin, sspec_out, ws_out);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_zero_u32(uint32_t* p, uint32_t count, hipStream_t s) {
    if (count > 64) return hipErrorInvalidValue;
    k_zero_u32<<<1, 64, 0, s>>>(p, count);
    return hipGetLastError();
}

hipError_t launch_transpose(const uint32_t* bits, uint32_t* tb, uint32_t* count, const Geom& g,
                            int is_seed, hipStream_t s) {
    switch (g.D) {
    case 1: return launch_transpose_t<1>(bits, tb, count, g, is_seed, s);
    case 2: return launch_transpose_t<2>(bits, tb, count, g, is_seed, s);
    case 4: return launch_transpose_t<4>(bits, tb, count, g, is_seed, s);
    case 8: return launch_transpose_t<8>(bits, tb, count, g, is_seed, s);
    case 16: return launch_transpose_t<16>(bits, tb, count, g, is_seed, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fwd_cols(const uint32_t* tb, float2* Y, const Geom& g, const Tables& t, hipStream_t s) {
    switch (g.lg1) {
    case 1: return launch_fwd_cols_t<1, 1>(tb, Y, g, t, s);
    case 2: return launch_fwd_cols_t<2, 1>(tb, Y, g, t, s);
    case 3: return launch_fwd_cols_t<3, 1>(tb, Y, g, t, s);
    case 4: return launch_fwd_cols_t<4, 1>(tb, Y, g, t, s);
    case 5: return launch_fwd_cols_t<5, 1>(tb, Y, g, t, s);
    case 6: return launch_fwd_cols_t<6, 1>(tb, Y, g, t, s);
    case 7: return launch_fwd_cols_t<7, 1>(tb, Y, g, t, s);
    case 8: return launch_fwd_cols_t<8, 1>(tb, Y, g, t, s);
    case 9: return launch_fwd_cols_t<9, 1>(tb, Y, g, t, s);
    case 10: return launch_fwd_cols_t<9, 2>(tb, Y, g, t, s);
    case 11: return launch_fwd_cols_t<9, 4>(tb, Y, g, t, s);
    case 12: return launch_fwd_cols_t<9, 8>(tb, Y, g, t, s);
    case 13: return launch_fwd_cols_t<9, 16>(tb, Y, g, t, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rows(int mode, float2* Yk, float2* S, const Geom& g, const Tables& t, hipStream_t s) {
    const uint32_t grid = g.N1 / 2u;
    switch (mode) {
    case ROWS_STATIC: k_rows<ROWS_STATIC><<<grid, ROWS_NT, 0, s>>>(Yk, S, g, t.tw14, t.lo, t.hi); break;
    case ROWS_SEEDPREP: k_rows<ROWS_SEEDPREP><<<grid, ROWS_NT, 0, s>>>(Yk, S, g, t.tw14, t.lo, t.hi); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_inv_cols(const float2* R, uint8_t* nib, const Geom& g, const uint32_t* wk,
                           const uint32_t* ws, const Tables& t, float* margin, hipStream_t s) {
    switch (g.lg1) {
    case 1: return launch_inv_cols_t<1>(R, nib, g, wk, ws, t, margin, s);
    case 2: return launch_inv_cols_t<2>(R, nib, g, wk, ws, t, margin, s);
    case 3: return launch_inv_cols_t<3>(R, nib, g, wk, ws, t, margin, s);
    case 4: return launch_inv_cols_t<4>(R, nib, g, wk, ws, t, margin, s);
    case 5: return launch_inv_cols_t<5>(R, nib, g, wk, ws, t, margin, s);
    case 6: return launch_inv_cols_t<6>(R, nib, g, wk, ws, t, margin, s);
    case 7: return launch_inv_cols_t<7>(R, nib, g, wk, ws, t, margin, s);
    case 8: return launch_inv_cols_t<8>(R, nib, g, wk, ws, t, margin, s);
    case 9: return launch_inv_cols_t<9>(R, nib, g, wk, ws, t, margin, s);
    case 10: return launch_inv_cols_t<10>(R, nib, g, wk, ws, t, margin, s);
    case 11: return launch_inv_cols_t<11>(R, nib, g, wk, ws, t, margin, s);
    case 12: return launch_inv_cols_t<12>(R, nib, g, wk, ws, t, margin, s);
    case 13: return launch_inv_cols_t<13>(R, nib, g, wk, ws, t, margin, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_pack(const uint8_t* nib, const uint32_t* key, uint32_t* out, const Geom& g, hipStream_t s) {
    const uint32_t words = g.out_bits >> 5;
    k_pack<<<(words + 255u) / 256u, 256, 0, s>>>(nib, key, out, g);
    return hipGetLastError();
}

hipError_t launch_small(const uint32_t* key, const uint32_t* seed, const float2* sspec,
                        const uint32_t* ws_ptr, uint32_t* out, const Geom& g, const Tables& t,
                        float* margin, float2* sspec_out, uint32_t* ws_out, hipStream_t s) {
    switch (g.lgN) {
    case 5: return launch_small_t<5>(key, seed, sspec, ws_ptr, out, g, t, margin, sspec_out, ws_out, s);
    case 6: return launch_small_t<6>(key, seed, sspec, ws_ptr, out, g, t, margin, sspec_out, ws_out, s);
    case 7: return launch_small_t<7>(key, seed, sspec, ws_ptr, out, g, t, margin, sspec_out, ws_out, s);
    case 8: return launch_small_t<8>(key, seed, sspec, ws_ptr, out, g, t, margin, sspec_out, ws_out, s);
    case 9: return launch_small_t<9>(key, seed, sspec, ws_ptr, out, g, t, margin, sspec_out, ws_out, s);
    case 10: return launch_small_t<10>(key, seed, sspec, ws_ptr, out, g, t, margin, sspec_out, ws_out, s);
    case 11: return launch_small_t<11>(key, seed, sspec, ws_ptr, out, g, t, margin, sspec_out, ws_out, s);
    case 12: return launch_small_t<12>(key, seed, sspec, ws_ptr, out, g, t, margin, sspec_out, ws_out, s);
    case 13: return launch_small_t<13>(key, seed, sspec, ws_ptr, out, g, t, margin, sspec_out, ws_out, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pa

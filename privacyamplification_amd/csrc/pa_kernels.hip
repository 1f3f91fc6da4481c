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
//   k_fwd_cols        pass A: decimated column FFTs of the +-1/2-encoded bits (R2C forward part);
//   k_fwd_cols9p      its persistent form for 512-point classes (the 2^27-bit block's shape)
//   k_rows            pass B: row FFTs, real split, DC zeroing (setFirstElementToZero, :408-418),
//                     complex product (ElementWiseProduct, :465-481), inverse split, inverse row FFTs
//   k_inv_cols        pass A': inverse column FFTs + round-mod-2 with the correction term
//                     (calculateCorrectionFloat :363-373, ToBinaryArray :739-783)
//   k_pack            output word assembly with the key-rest XOR and byte order
//                     (key2StartRest's key_rest :873-878, ToBinaryArray's bswap XOR :757-762)
//   k_small           the whole block in one workgroup for L <= 2^14
#include <hip/hip_runtime.h>

#include <cstdio>
#include <type_traits>

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
// Word widx (< L/32) of the seed sequence t: the n seed bits rotated left by seed_rot words and
// repeated with period n up to the transform length L.  The output window only reads t[0, n),
// where the repetition changes nothing; beyond n it keeps the +-1/2-encoded sequence free of a
// constant block.  A zero pad there would become a -1/2 block whose convolution with the key's
// pad swings by about h/4 across positions after the mean is removed, far beyond FP32's
// resolution at L = 2^24; the repeated seed leaves every position a random-walk sum of the same
// statistics, and every sum in the correction stays constant (ws counts all L positions).
__device__ __forceinline__ uint32_t seed_word(const uint32_t* __restrict__ seed, const Geom& g, uint32_t widx) {
    uint32_t s = widx + g.seed_rot;  // < 3 seed_words: L < 2n when L > n
    if (s >= g.seed_words) s -= g.seed_words;
    if (s >= g.seed_words) s -= g.seed_words;
    return seed[s];
}

// Popcounts are accumulated in POP_SLOTS partial counters (spread atomics); readers sum them.
__device__ __forceinline__ uint32_t pop_total(const uint32_t* __restrict__ c) {
    uint32_t t = 0;
#pragma unroll 8
    for (int i = 0; i < POP_SLOTS; ++i) t += c[i];
    return t;
}


// Batched launches (BATCH = true) take each block's geometry and buffers from bt[blockIdx.z].
template <bool BATCH>
__global__ __launch_bounds__(256) void k_zero_u32(uint32_t* p, uint32_t count, const BlockArgs* __restrict__ bt) {
    if constexpr (BATCH) p = bt[blockIdx.z].counts;
    for (uint32_t i = threadIdx.x; i < count; i += 256) p[i] = 0u;
}

// ------------------------------------------------------------------ staging copies ----
// One launch moves every staging range of a batch: workgroup (i, y) copies chunks y, y + grid.y,
// ... of COPY_CHUNK bytes of range d[i].  Host ranges are pinned, device-mapped memory that the
// kernel reads or writes over the bus, so a batch costs one launch instead of one DMA command
// per key, seed and output.  The descriptors themselves may live in mapped host memory.
constexpr uint32_t COPY_NT = 256;
constexpr uint64_t COPY_CHUNK = 64u * 1024u;
__global__ __launch_bounds__(COPY_NT) void k_copies(const CopyDesc* __restrict__ d) {
    const uint8_t* src = d[blockIdx.x].src;
    uint8_t* dst = d[blockIdx.x].dst;
    const uint64_t bytes = d[blockIdx.x].bytes;
    const uint64_t chunks = (bytes + COPY_CHUNK - 1) / COPY_CHUNK;
    const uint32_t al = (uint32_t)(((uintptr_t)src | (uintptr_t)dst) & 15u);
    for (uint64_t ch = blockIdx.y; ch < chunks; ch += gridDim.y) {
        const uint64_t lo = ch * COPY_CHUNK;
        const uint32_t len = (uint32_t)(bytes - lo < COPY_CHUNK ? bytes - lo : COPY_CHUNK);
        uint32_t body = 0;
        if (al == 0) {  // 16-byte words, four loads in flight per thread
            const uint4* s4 = reinterpret_cast<const uint4*>(src + lo);
            uint4* d4 = reinterpret_cast<uint4*>(dst + lo);
            const uint32_t n16 = len / 16;
            for (uint32_t i = threadIdx.x; i < n16; i += 4 * COPY_NT) {
                uint4 v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (i + k * COPY_NT < n16) v[k] = s4[i + k * COPY_NT];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (i + k * COPY_NT < n16) d4[i + k * COPY_NT] = v[k];
            }
            body = n16 * 16;
        } else if ((al & 3u) == 0) {
            const uint32_t* s1 = reinterpret_cast<const uint32_t*>(src + lo);
            uint32_t* d1 = reinterpret_cast<uint32_t*>(dst + lo);
            const uint32_t n4 = len / 4;
            for (uint32_t i = threadIdx.x; i < n4; i += COPY_NT) d1[i] = s1[i];
            body = n4 * 4;
        }
        for (uint32_t i = body + threadIdx.x; i < len; i += COPY_NT) dst[lo + i] = src[lo + i];
    }
}

// ------------------------------------------------------------------ bit transpose ----
// One round of the recursive 32 x 32 bit-matrix transpose (rows A[k], column 0 = the MSB): swap
// the off-diagonal J x J blocks of every 2J x 2J block.  Five rounds J = 16, 8, 4, 2, 1 transpose.
template <int J>
__device__ __forceinline__ void bit_transpose_round(uint32_t (&A)[32]) {
    constexpr uint32_t M = J == 16 ? 0x0000FFFFu : J == 8 ? 0x00FF00FFu : J == 4 ? 0x0F0F0F0Fu
                         : J == 2 ? 0x33333333u : 0x55555555u;
#pragma unroll
    for (int b = 0; b < 32; b += 2 * J)
#pragma unroll
        for (int r = 0; r < J; ++r) {
            const uint32_t t = (A[b + r] ^ (A[b + r + J] >> J)) & M;
            A[b + r] ^= t;
            A[b + r + J] ^= t << J;
        }
}

// tb[u*N2 + t2] = E | O << 16 with E bit v = x[2(t2 + N2(u + M v))], O bit v = the odd bit.
// Workgroup (u, part): thread k covers t2 in [part*TB_SPAN + 32k, +32): for each class v it reads
// the two words holding those 64 bits (coalesced across the workgroup), transposes the D x 32
// bit-pairs in registers, and the workgroup writes its TB_SPAN outputs through LDS.  Popcounts
// are reduced per workgroup into one of POP_SLOTS counters.
template <int D, bool BATCH>
__global__ __launch_bounds__(TB_NT) void k_transpose_bits(const uint32_t* __restrict__ bits,
                                                          uint32_t* __restrict__ tb,
                                                          uint32_t* __restrict__ count, Geom g,
                                                          int is_seed, const BlockArgs* __restrict__ bt,
                                                          const uint32_t* __restrict__ bits2,
                                                          uint32_t* __restrict__ tb2) {
    __shared__ uint32_t stage[TB_NT * 33];
    __shared__ uint32_t red[TB_NT / 64];
    if (!BATCH && blockIdx.y == 1) {  // key and seed in one launch: grid.y selects the seed
        bits = bits2;
        tb = tb2;
        count += POP_SLOTS;
        is_seed = 1;
    }
    if constexpr (BATCH) {
        const BlockArgs& a = bt[blockIdx.z];
        if (is_seed && a.seed == nullptr) return;  // resident seed: nothing to transpose
        g = a.g;
        bits = is_seed ? a.seed : a.key;
        tb = is_seed ? a.tbs : a.tbk;
        count = a.counts + (is_seed ? POP_SLOTS : 0);
    }
    constexpr int PARTS_LOG = ROW_LOG - TB_SPAN_LOG;
    const uint32_t u = blockIdx.x >> PARTS_LOG, part = blockIdx.x & ((1u << PARTS_LOG) - 1u);
    const uint32_t t0 = (part << TB_SPAN_LOG) + 32u * threadIdx.x;
    uint32_t w0[D], w1[D];
    uint32_t pc = 0;
#pragma unroll
    for (int v = 0; v < D; ++v) {
        const uint32_t widx = (2u * (t0 + g.N2 * (u + g.M * (uint32_t)v))) >> 5;  // even
        w0[v] = is_seed ? seed_word(bits, g, widx) : key_word(bits, g, widx);
        w1[v] = is_seed ? seed_word(bits, g, widx + 1u) : key_word(bits, g, widx + 1u);
        pc += (uint32_t)(__popc(w0[v]) + __popc(w1[v]));
    }
    uint32_t* srow = stage + threadIdx.x * 33;
    if constexpr (D >= 8) {
        // Rows A[15 - v] = w1[v], A[31 - v] = w0[v] (zero for v >= D), transposed: A[c] then holds
        // bit 31 - c of w0[v] at bit v and of w1[v] at bit v + 16, so output i (E bits from the
        // pair's even bit 2i, MSB first, O bits from the odd one) is two half-words.
        uint32_t A[32];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            A[15 - v] = v < D ? w1[v < D ? v : 0] : 0u;
            A[31 - v] = v < D ? w0[v < D ? v : 0] : 0u;
        }
        bit_transpose_round<16>(A);
        bit_transpose_round<8>(A);
        bit_transpose_round<4>(A);
        bit_transpose_round<2>(A);
        bit_transpose_round<1>(A);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            srow[i] = (A[2 * i] & 0xFFFFu) | (A[2 * i + 1] << 16);
            srow[16 + i] = (A[2 * i] >> 16) | (A[2 * i + 1] & 0xFFFF0000u);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            const int sh = 31 - 2 * (i & 15);  // even bit 2i of the 64-bit pair, MSB first
            uint32_t E = 0, O = 0;
#pragma unroll
            for (int v = 0; v < D; ++v) {
                const uint32_t w = i < 16 ? w0[v] : w1[v];
                E |= ((w >> sh) & 1u) << v;
                O |= ((w >> (sh - 1)) & 1u) << v;
            }
            srow[i] = E | (O << 16);
        }
    }
    for (int off = 32; off > 0; off >>= 1) pc += __shfl_xor(pc, off);
    if ((threadIdx.x & 63u) == 0u) red[threadIdx.x >> 6] = pc;
    __syncthreads();
    uint32_t* out = tb + (size_t)u * g.N2 + (part << TB_SPAN_LOG);
    for (int e = threadIdx.x; e < (1 << TB_SPAN_LOG); e += TB_NT) out[e] = stage[(e >> 5) * 33 + (e & 31)];
    if (threadIdx.x == 0) {
        uint32_t t = 0;
#pragma unroll
        for (int k = 0; k < TB_NT / 64; ++k) t += red[k];
        if (t) atomicAdd(count + (blockIdx.x % POP_SLOTS), t);
    }
}

// ------------------------------------------------------------------ pass A ----
// Workgroup (column group grp of CPW columns, class c): computes Y[k1][t2] for k1 = c + D*k1',
// t2 in the group:  Y = w_N^{t2 k1} * FFT_M( w_N1^{u c} * sum_v z[t2 + N2(u + M v)] w_D^{v c} ).
// The class sum over v is one or two 256-entry table lookups per operand half (bits of tb).  The
// class values go to LDS and the M-point FFT runs there.  M = 512 (N1 = 512 D, the 2^27-bit
// block's shape) runs in k_fwd_cols9p instead.
template <int LGM, int D, bool BATCH, int CPW = fwd_cols_width(LGM)>
__global__ __launch_bounds__(fwd_cols_threads(LGM)) void k_fwd_cols(const uint32_t* __restrict__ tb,
                                                     float2* __restrict__ Y, Geom g,
                                                     const float2* __restrict__ tw14,
                                                     const float2* __restrict__ twlo,
                                                     const float2* __restrict__ twhi,
                                                     const float2* __restrict__ cls,
                                                     const BlockArgs* __restrict__ bt, int is_seed,
                                                     const uint32_t* __restrict__ tb2,
                                                     float2* __restrict__ Y2) {
    if (!BATCH && blockIdx.y == 1) {  // key and seed in one launch: grid.y selects the seed
        tb = tb2;
        Y = Y2;
    }
    if constexpr (BATCH) {
        const BlockArgs& a = bt[blockIdx.z];
        if (is_seed && a.seed == nullptr) return;
        g = a.g;
        tb = is_seed ? a.tbs : a.tbk;
        Y = is_seed ? a.Ys : a.Yk;
    }
    constexpr int M = 1 << LGM;
    constexpr int NT = fwd_cols_threads(LGM);
    constexpr int SEQS = seq_stride(M);
    static_assert(LGM != 9, "512-point classes run in k_fwd_cols9p");
    __shared__ float2 buf[CPW * SEQS];
    __shared__ float2 twl[TWL_SIZE];
    constexpr int BW = cls_bits(D);         // bits per table: v in [BW q, BW q + BW)
    constexpr int NB = cls_tables(D);       // tables per 16-bit half
    constexpr uint32_t MASK = (1u << BW) - 1u;
    __shared__ float2 TB[NB][1 << BW];      // TB[q][byte] = sum_b bit_b(byte) w_D^{(BW q + b) c}
    __shared__ float2 TU[M];                // w_N1^{u c}
    const uint32_t logical = xcd_logical(blockIdx.x, gridDim.x);
    const uint32_t c = logical % D, grp = logical / D;  // columns [CPW grp, CPW grp + CPW)
    load_twl(twl, tw14, NT);
    {
        const float2* src = cls + cls_table_offset(D) + (size_t)c * (NB << BW);
        for (int e = threadIdx.x; e < NB << BW; e += NT) TB[e >> BW][e & (int)MASK] = src[e];
    }
    // w_N1^{u c}: from the 2^14 table while N1 <= 2^14, else from the length-L two-level table
    if (g.N1 <= (uint32_t)TW14) {
        const uint32_t step_n1 = TW14 / g.N1;
        for (int u = threadIdx.x; u < M; u += NT) TU[u] = tw14[((uint32_t)u * c) * step_n1];
    } else {
        const uint32_t step_l = g.L / g.N1;
        for (int u = threadIdx.x; u < M; u += NT) TU[u] = tw_two_level(twlo, twhi, ((uint32_t)u * c) * step_l);
    }
    // the element z(col, u) of this class (both halves of tb: even and odd bits)
    auto zval = [&](uint32_t t, int u, auto is_c0) {
        float2 sE = TB[0][t & MASK], sO = TB[0][(t >> 16) & MASK];
#pragma unroll
        for (int q = 1; q < NB; ++q) {
            sE = cadd(sE, TB[q][(t >> (BW * q)) & MASK]);
            sO = cadd(sO, TB[q][(t >> (16 + BW * q)) & MASK]);
        }
        float2 z = make_float2(sE.x - sO.y, sE.y + sO.x);
        if constexpr (decltype(is_c0)::value) {
            z.x -= 0.5f * (float)D;
            z.y -= 0.5f * (float)D;
        } else {
            z = cmul(z, TU[u]);
        }
        return z;
    };
    constexpr int IPT = (CPW * M + NT - 1) / NT;
    uint32_t tbv[IPT];
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
        const int e = threadIdx.x + i * NT;
        tbv[i] = (e < CPW * M) ? tb[(uint32_t)(e / CPW) * g.N2 + grp * CPW + (uint32_t)(e % CPW)] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
        const int e = threadIdx.x + i * NT;
        if (e < CPW * M)
            buf[(e % CPW) * SEQS + padi(e / CPW)] =
                c == 0 ? zval(tbv[i], e / CPW, std::true_type{}) : zval(tbv[i], e / CPW, std::false_type{});
    }
    __syncthreads();
    lds_fft<LGM, CPW, SEQS, NT>(buf, twl);
    // Output e = tid + NT i is column tid mod CPW, k1' = tid / CPW + RPI i: its twiddle
    // w_L^{2 t2 k1} is base step^i, re-anchored from the table every 8 steps (a gather per point
    // cost more than the transform at N1 >= 2^14)
    static_assert(NT % CPW == 0, "a thread keeps one column");
    constexpr int RPI = NT / CPW;
    const int col = threadIdx.x % CPW, jt = threadIdx.x / CPW;
    const uint32_t t2 = grp * CPW + (uint32_t)col;
    const uint64_t L1m = g.L - 1u;
    auto tw_at = [&](uint32_t k1p) {
        return tw_two_level(twlo, twhi, (uint32_t)((2ull * t2 * (c + (uint64_t)D * k1p)) & L1m));
    };
    const float2 step = tw_two_level(twlo, twhi, (uint32_t)((2ull * t2 * (uint64_t)D * RPI) & L1m));
    float2 tw = make_float2(1.f, 0.f);
    for (int i = 0; i < (CPW * M + NT - 1) / NT; ++i) {
        const uint32_t k1p = (uint32_t)jt + (uint32_t)RPI * (uint32_t)i;
        if (k1p >= (uint32_t)M) break;  // CPW M < NT: the upper threads have no output
        tw = (i & 7) == 0 ? tw_at(k1p) : cmul(tw, step);
        const uint32_t k1 = c + (uint32_t)D * k1p;
        st_nt(&Y[(size_t)k1 * g.N2 + t2], cmul(buf[col * SEQS + padi((int)k1p)], tw));
    }
}

// Raw buffer access (vector memory instructions): one per-lane byte offset plus a uniform
// scalar offset, so a run of accesses at fixed strides keeps a single address register.
typedef unsigned int pa_u2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ void buf_store_nt(float2 v, __amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
    const pa_u2v x = {__float_as_uint(v.x), __float_as_uint(v.y)};
    __builtin_amdgcn_raw_buffer_store_b64(x, rs, (int)voff, (int)soff, 2);  // aux 2: nt
}
__device__ __forceinline__ float2 buf_load_f2(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
    const pa_u2v x = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, (int)soff, 0);
    return make_float2(__uint_as_float(x.x), __uint_as_float(x.y));
}
__device__ __forceinline__ uint32_t buf_load_u32(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
    return __builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, (int)soff, 0);
}

// Pieces of the persistent pass A (k_fwd_cols9p).  Thread (col = t mod 16, jt = t / 16).
// Bit words of column group grp: stage-1 butterflies j = jt + 16 i, inputs u = j + 32 r.
__device__ __forceinline__ void fwd9_load_tv(uint32_t (&tv)[2][16], __amdgpu_buffer_rsrc_t rtb, uint32_t N2,
                                             uint32_t t2, int jt) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const uint32_t voff = ((uint32_t)(jt + 16 * i) * N2 + t2) * 4u;
#pragma unroll
        for (int r = 0; r < 16; ++r) tv[i][r] = buf_load_u32(rtb, voff, (uint32_t)(32 * r) * N2 * 4u);
    }
}
// Table entries of the output twiddles w_N^{t2 k1}, k1 = c + D (jt + 16 r) = base step^r, with
// anchors w_N^{t2 16 D 8 (a+1)}: (lo, hi) pairs of base, step, anchor 0..2.  The products are
// formed where they are used, so that gathering them never waits on other loads.
template <int D>
__device__ __forceinline__ void fwd9_load_tw(float2 (&tw)[10], const float2* __restrict__ lo,
                                             const float2* __restrict__ hi, uint32_t L, uint32_t c, uint32_t t2,
                                             int jt) {
    const uint64_t T = t2;
    uint32_t x[5];
    x[0] = (uint32_t)((2ull * T * (c + (uint64_t)D * (uint32_t)jt)) & (L - 1u));
    x[1] = (uint32_t)((2ull * T * 16ull * D) & (L - 1u));
#pragma unroll
    for (int a = 0; a < 3; ++a) x[2 + a] = (uint32_t)((2ull * T * 16ull * D * (8 * (a + 1))) & (L - 1u));
#pragma unroll
    for (int e = 0; e < 5; ++e) {
        tw[2 * e] = lo[x[e] & (TW14 - 1)];
        tw[2 * e + 1] = hi[x[e] >> TW14_LOG];
    }
}
template <int D, bool C0>
__device__ __forceinline__ void fwd9_stage1(const uint32_t (&tv)[2][16], float2* buf,
                                            const float2 (*TB)[1 << cls_bits(D)], const float2* TU, int col, int jt) {
    constexpr int SEQS = seq_stride(512);
    constexpr int BW = cls_bits(D);
    constexpr int NB = cls_tables(D);
    constexpr uint32_t MASK = (1u << BW) - 1u;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int j = jt + 16 * i;
        float2 v[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const uint32_t t = tv[i][r];
            float2 sE = TB[0][t & MASK], sO = TB[0][(t >> 16) & MASK];
#pragma unroll
            for (int q = 1; q < NB; ++q) {
                sE = cadd(sE, TB[q][(t >> (BW * q)) & MASK]);
                sO = cadd(sO, TB[q][(t >> (16 + BW * q)) & MASK]);
            }
            float2 z = make_float2(sE.x - sO.y, sE.y + sO.x);
            if constexpr (C0) {
                z.x -= 0.5f * (float)D;
                z.y -= 0.5f * (float)D;
            } else {
                z = cmul(z, TU[j + 32 * r]);
            }
            v[r] = z;
        }
        stage_first_to_lds<16>(v, buf + col * SEQS, j);
    }
}
// tw2: the radix-32 stage's twiddle table, w_512^(jt r) at tw2[jt * 33 + r] (fill_stage_table<32, 16>).
template <int D>
__device__ __forceinline__ void fwd9_stage2(const float2* buf, const float2* tw2, const float2 (&tw)[10],
                                            __amdgpu_buffer_rsrc_t ry, uint32_t N2, uint32_t c, uint32_t t2,
                                            int col, int jt) {
    constexpr int SEQS = seq_stride(512);
    float2 F[32];
    stage_last_from_lds_tab<32, 512>(F, buf + col * SEQS, jt, tw2 + jt * 33);
    const float2 base = cmul(tw[0], tw[1]), step = cmul(tw[2], tw[3]);
    const uint32_t voff = ((c + (uint32_t)D * (uint32_t)jt) * N2 + t2) * 8u;
    float2 w = base;
#pragma unroll
    for (int r = 0; r < 32; ++r) {
        if ((r & 7) == 0 && r) w = cmul(base, cmul(tw[2 * (1 + r / 8)], tw[2 * (1 + r / 8) + 1]));
        buf_store_nt(cmul(F[r], w), ry, voff, (uint32_t)(16 * D * r) * N2 * 8u);
        w = cmul(w, step);
    }
}
template <int D, bool C0>
__device__ __forceinline__ void fwd9_walk(__amdgpu_buffer_rsrc_t rtb, __amdgpu_buffer_rsrc_t ry, const Geom& g,
                                          const float2* __restrict__ twlo, const float2* __restrict__ twhi,
                                          float2* buf, const float2* tw2, const float2 (*TB)[1 << cls_bits(D)],
                                          const float2* TU, uint32_t c, uint32_t grp0, uint32_t per) {
    const int col = threadIdx.x & 15;
    const int jt = threadIdx.x >> 4;
    const uint32_t N2 = g.N2;
    uint32_t tv[2][16];
    float2 tw[10], twn[10];
    // group grp0, peeled: every entry into the loop below then has the same memory-operation
    // history (next group's loads, then this group's stores), so the compiler's waits on the
    // prefetched registers count past the stores instead of draining them
    fwd9_load_tv(tv, rtb, N2, grp0 * 16u + (uint32_t)col, jt);
    fwd9_load_tw<D>(tw, twlo, twhi, g.L, c, grp0 * 16u + (uint32_t)col, jt);
    __syncthreads();  // tables
    fwd9_stage1<D, C0>(tv, buf, TB, TU, col, jt);
    __syncthreads();
    if (per > 1) {
        fwd9_load_tv(tv, rtb, N2, (grp0 + 1u) * 16u + (uint32_t)col, jt);
        fwd9_load_tw<D>(twn, twlo, twhi, g.L, c, (grp0 + 1u) * 16u + (uint32_t)col, jt);
    }
    fwd9_stage2<D>(buf, tw2, tw, ry, N2, c, grp0 * 16u + (uint32_t)col, col, jt);
    for (uint32_t it = 1; it < per; ++it) {
        const uint32_t grp = grp0 + it;
#pragma unroll
        for (int e = 0; e < 10; ++e) tw[e] = twn[e];
        __syncthreads();  // stage-2 reads of the previous group are done
        fwd9_stage1<D, C0>(tv, buf, TB, TU, col, jt);
        __syncthreads();
        if (it + 1 < per) {
            fwd9_load_tv(tv, rtb, N2, (grp + 1u) * 16u + (uint32_t)col, jt);
            fwd9_load_tw<D>(twn, twlo, twhi, g.L, c, (grp + 1u) * 16u + (uint32_t)col, jt);
        }
        fwd9_stage2<D>(buf, tw2, tw, ry, N2, c, grp * 16u + (uint32_t)col, col, jt);
    }
}

// Persistent pass A for M = 512 (the fused 16 x 32 plan of k_fwd_cols): workgroup L of G keeps one
// class c = L mod D and walks the column groups [k per, k per + per), k = L / D, so its class
// tables are loaded once, and each group's stores drain while the next group is transformed (its
// loads were issued before those stores).  The D workgroups of one k sit on one XCD and walk the
// same groups in step, so each bit word is fetched into L2 once.
template <int D, bool BATCH>
__global__ __launch_bounds__(FWD_NT, 2) void k_fwd_cols9p(const uint32_t* __restrict__ tb,
                                                          float2* __restrict__ Y, Geom g,
                                                          const float2* __restrict__ tw14,
                                                          const float2* __restrict__ twlo,
                                                          const float2* __restrict__ twhi,
                                                          const float2* __restrict__ cls,
                                                          const uint32_t* __restrict__ tb2,
                                                          float2* __restrict__ Y2, uint32_t per,
                                                          const BlockArgs* __restrict__ bt, int is_seed) {
    if constexpr (BATCH) {  // grid.z = block of a batch group (one transform length)
        const BlockArgs& a = bt[blockIdx.z];
        if (is_seed && a.seed == nullptr) return;  // resident seed: nothing to transform
        g = a.g;
        tb = is_seed ? a.tbs : a.tbk;
        Y = is_seed ? a.Ys : a.Yk;
    } else if (blockIdx.y == 1) {  // key and seed in one launch: grid.y selects the seed
        tb = tb2;
        Y = Y2;
    }
    constexpr int M = 512;
    constexpr int NT = FWD_NT;
    constexpr int SEQS = seq_stride(M);
    static_assert(NT == 256, "fused pass A expects 256 threads");
    __shared__ float2 buf[16 * SEQS];
    __shared__ float2 tw2[16 * 33];  // radix-32 stage twiddles w_512^(jt r), rows padded to 33
    constexpr int BW = cls_bits(D);
    constexpr int NB = cls_tables(D);
    constexpr uint32_t MASK = (1u << BW) - 1u;
    __shared__ float2 TB[NB][1 << BW];
    __shared__ float2 TU[M];
    const uint32_t logical = xcd_logical(blockIdx.x, gridDim.x);
    const uint32_t c = logical % D, grp0 = (logical / D) * per;
    const __amdgpu_buffer_rsrc_t rtb = make_rsrc(tb, (uint32_t)M * g.N2 * 4u);
    const __amdgpu_buffer_rsrc_t ry = make_rsrc(Y, g.N1 * g.N2 * 8u);
    fill_stage_table<32, 16>(tw2, tw14, NT);
    {
        const float2* src = cls + cls_table_offset(D) + (size_t)c * (NB << BW);
        for (int e = threadIdx.x; e < NB << BW; e += NT) TB[e >> BW][e & (int)MASK] = src[e];
    }
    if (g.N1 <= (uint32_t)TW14) {
        const uint32_t step_n1 = TW14 / g.N1;
        for (int u = threadIdx.x; u < M; u += NT) TU[u] = tw14[((uint32_t)u * c) * step_n1];
    } else {
        const uint32_t step_l = g.L / g.N1;
        for (int u = threadIdx.x; u < M; u += NT) TU[u] = tw_two_level(twlo, twhi, ((uint32_t)u * c) * step_l);
    }
    if (c == 0) fwd9_walk<D, true>(rtb, ry, g, twlo, twhi, buf, tw2, TB, TU, c, grp0, per);
    else fwd9_walk<D, false>(rtb, ry, g, twlo, twhi, buf, tw2, TB, TU, c, grp0, per);
}

// ------------------------------------------------------------------ pass B ----
// Item w (one row pair) owns rows (w, N1-w), or (0, N1/2) for w = 0: the rows whose frequencies
// pair k <-> N-k in the real split (k = row + N1*k2).  Row transforms have N2 = 8192 points:
//   forward plan 32 x 16 x 16: stage 1 straight from global memory, stage 3 into registers;
//   inverse plan 16 x 16 x 32: stage 1 from registers, stage 3 straight to global memory.
// After forward stage 3 a thread holds Z at k2 = jA + 512 r (row sA) and k2 = jB + 512 r (row sB)
// for r < 16, chosen so that the partners of the first are the second read backwards:
//   w != 0:                     (row0, t)      and (row1, 511-t)        pairs (A r, B 15-r)
//   w == 0, 0 < t < 256:        (row0, t)      and (row0, 512-t)        pairs (A r, B 15-r)
//   w == 0, t >= 256:           (row1, t-256)  and (row1, 767-t)        pairs (A r, B 15-r)
//   w == 0, t == 0:             (row0, 0)      and (row0, 256)          pairs (A r, A 16-r) with
//                               k = 0 (DC zeroed, Nyquist carried) and k = N/2 self-paired,
//                               and (B r, B 15-r)
// so the real split, product and inverse split need no data exchange, and their results are
// exactly the inputs of inverse stage 1.
//   ROWS_SEEDPREP: seed rows (pass-A output) -> split seed spectrum S, in place (+ S[N] at index N).
//   ROWS_STATIC:   key rows + resident S -> R'[k1][t2] = w_N^{t2 k1} FFT_N2(conj Q)[t2], in place,
//                  or into pass A''s tiles (Rt) when rt_tiled(lg1).
//   ROWS_DYN:      key rows + seed rows; S is formed first and held in registers.
// Persistent: each workgroup walks the items blockIdx.x, +gridDim.x, ...; the global inputs of
// the next transform are requested while the current one runs (register prefetch).
struct RowMap {
    uint32_t row0, row1, rowA, rowB;
    int sA, jA, sB, jB;
    bool special;
};
__device__ __forceinline__ RowMap row_map(uint32_t w, uint32_t N1, int tid) {
    RowMap m;
    const bool w0 = (w == 0);
    m.row0 = w;
    m.row1 = w0 ? (N1 >> 1) : (N1 - w);
    m.special = w0 && tid == 0;
    if (!w0) { m.sA = 0; m.jA = tid; m.sB = 1; m.jB = 511 - tid; }
    else if (tid == 0) { m.sA = 0; m.jA = 0; m.sB = 0; m.jB = 256; }
    else if (tid < 256) { m.sA = 0; m.jA = tid; m.sB = 0; m.jB = 512 - tid; }
    else { m.sA = 1; m.jA = tid - 256; m.sB = 1; m.jB = 767 - tid; }
    m.rowA = m.sA ? m.row1 : m.row0;
    m.rowB = m.sB ? m.row1 : m.row0;
    return m;
}

template <int MODE, bool BATCH>
__global__ __launch_bounds__(ROWS_NT) void k_rows(float2* __restrict__ Yk, float2* __restrict__ Sst,
                                                  float2* __restrict__ Rt, Geom g, const float2* __restrict__ tw14,
                                                  const float2* __restrict__ twlo,
                                                  const float2* __restrict__ twhi,
                                                  const BlockArgs* __restrict__ bt) {
    if constexpr (BATCH) {
        const BlockArgs& a = bt[blockIdx.z];
        g = a.g;
        Yk = a.Yk;
        Sst = (MODE == ROWS_STATIC) ? const_cast<float2*>(a.sspec) : a.Ys;
        Rt = a.Rt;
    }
    constexpr int N2 = 1 << ROW_LOG, NT = ROWS_NT, SEQS = seq_stride(N2);
    static_assert(N2 == 8192 && NT == 512, "row pass is written for 8192-point rows, 512 threads");
    static_assert(2 * N2 == TW14, "row split twiddle assumes N2 = 2^13");
    __shared__ float2 buf[2 * SEQS];
    __shared__ float2 twl[TWL_SIZE];
    __shared__ float2 ttf[32 * 17];  // forward middle stage (radix 16, span 32)
    __shared__ float2 tti[16 * 17];  // inverse middle stage (radix 16, span 16)
    load_twl(twl, tw14, NT);
    fill_stage_table<16, 32>(ttf, tw14, NT);
    fill_stage_table<16, 16>(tti, tw14, NT);
    const uint32_t N1 = g.N1, items = N1 >> 1;
    const int tid = threadIdx.x;
    // stage-1 butterfly of this thread: row (t >> 8), j = t & 255, inputs j + 256 r
    const int s1seq = tid >> 8, s1j = tid & 255;

    // Buffer loads: one 32-bit offset register per thread.  A 64-bit row pointer here was kept
    // spilled, and its reload before the next item's request waited for the output stores.
    auto s1_load = [&](const float2* __restrict__ src, uint32_t w, float2 (&v)[32]) {
        const uint32_t row = s1seq ? ((w == 0) ? (N1 >> 1) : (N1 - w)) : w;
        const __amdgpu_buffer_rsrc_t rs = make_rsrc(src, N1 * N2 * 8u);
        const uint32_t vo = (row * N2 + (uint32_t)s1j) * 8u;
#pragma unroll
        for (int r = 0; r < 32; ++r) v[r] = buf_load_f2(rs, vo, 2048u * r);
    };
    // forward stages 1 (inputs in v) .. 3; after stage 1 has consumed v, refill(v) requests the
    // next transform's inputs into the same registers; ends with this thread's stage-3 outputs
    auto forward = [&](float2 (&v)[32], const RowMap& m, float2 (&zA)[16], float2 (&zB)[16],
                       auto&& refill) {
        stage_first_to_lds<32>(v, buf + s1seq * SEQS, s1j);
        refill(v);
        __syncthreads();
        stockham_stage<16, 32, ROW_LOG, 2, SEQS, NT, true>(buf, ttf);
        stage_last_from_lds<16, N2>(zA, buf + m.sA * SEQS, m.jA, twl);
        stage_last_from_lds<16, N2>(zB, buf + m.sB * SEQS, m.jB, twl);
    };

    uint32_t w = blockIdx.x;
    if (w >= items) return;
    // vmcnt retires in issue order, so a prefetch only overlaps work if no later-issued global
    // load is waited on before the prefetched data is used: the small gathers of the next item
    // (row twiddles) are issued and consumed before its prefetch.
    RowMap m = row_map(w, N1, tid);
    float2 wrA = tw_two_level(twlo, twhi, m.rowA), wrB = tw_two_level(twlo, twhi, m.rowB);
    float2 pre[32];  // prefetched stage-1 inputs of the next transform
    s1_load(MODE == ROWS_STATIC ? Yk : Sst, w, pre);
    __syncthreads();  // twl

    while (true) {
        // split twiddles w_L^k at k = row + N1 (j + 512 r): w_L^row w_2^14^j w_32^r
        const float2 baseA = cmul(wrA, tw_lds(twl, m.jA)), baseB = cmul(wrB, tw_lds(twl, m.jB));
        auto wkA = [&](int r) { return w32pow(opaque_f2(baseA), r); };
        auto wkB = [&](int r) { return w32pow(opaque_f2(baseB), r); };
        const uint32_t wn = w + gridDim.x;
        const bool more = wn < items;
        RowMap mn = m;
        float2 wrAn = wrA, wrBn = wrB;

        float2 SA[16], SB[16];
        float2 Sn = make_float2(0.f, 0.f);
        if constexpr (MODE == ROWS_SEEDPREP || MODE == ROWS_DYN) {
            float2 zA[16], zB[16];
            // next request: the key rows of this item (DYN) or the seed rows of the next item
            forward(pre, m, zA, zB, [&](float2 (&v)[32]) {
                if (MODE == ROWS_DYN) s1_load(Yk, w, v);
                else if (more) s1_load(Sst, wn, v);
            });
            if (!m.special) {
#pragma unroll
                for (int r = 0; r < 16; ++r) r2c_split_pair(zA[r], zB[15 - r], wkA(r), SA[r], SB[15 - r]);
            } else {
                Sn = r2c_nyquist(zA[0]);
                SA[0] = make_float2(0.f, 0.f);
                SA[8] = r2c_split(zA[8], zA[8], wkA(8));
#pragma unroll
                for (int r = 1; r < 8; ++r) {
                    SA[r] = r2c_split(zA[r], zA[16 - r], wkA(r));
                    SA[16 - r] = r2c_split(zA[16 - r], zA[r], wkA(16 - r));
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) SB[r] = r2c_split(zB[r], zB[15 - r], wkB(r));
            }
            __syncthreads();  // stage-3 reads done before LDS is reused
            if constexpr (MODE == ROWS_SEEDPREP) {
                // in place: every thread of the workgroup has consumed these rows
                if (more) {
                    mn = row_map(wn, N1, tid);
                    wrAn = tw_two_level(twlo, twhi, mn.rowA);
                    wrBn = tw_two_level(twlo, twhi, mn.rowB);
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    st_nt(&Sst[(size_t)m.rowA * N2 + m.jA + 512 * r], SA[r]);
                    st_nt(&Sst[(size_t)m.rowB * N2 + m.jB + 512 * r], SB[r]);
                }
                if (m.special) Sst[(size_t)N1 * N2] = Sn;
                if (!more) break;
                w = wn;
                m = mn;
                wrA = wrAn;
                wrB = wrBn;
                continue;
            }
        }

        // key rows: forward, split, product, inverse split -> conj(Q) at the same positions
        float2 qA[16], qB[16];
        {
            float2 zA[16], zB[16];
            forward(pre, m, zA, zB, [&](float2 (&)[32]) {
                if constexpr (MODE == ROWS_STATIC) {
                    // the resident spectrum of this item is requested now, used after the FFT
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        SA[r] = Sst[(size_t)m.rowA * N2 + m.jA + 512 * r];
                        SB[r] = Sst[(size_t)m.rowB * N2 + m.jB + 512 * r];
                    }
                    if (m.special) Sn = Sst[(size_t)N1 * N2];
                }
            });
            auto unit = [&](float2 za, float2 zb, float2 sa, float2 sb, float2 wa, float2 wb, float2& qa,
                            float2& qb) {
                const float2 pa = cmul(r2c_split(za, zb, wa), sa);
                const float2 pb = cmul(r2c_split(zb, za, wb), sb);
                qa = conjf2(c2r_split(pa, pb, wa));
                qb = conjf2(c2r_split(pb, pa, wb));
            };
            if (!m.special) {
                // partner pairs (A r, B 15-r): k and N-k, one twiddle for both
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const float2 w = wkA(r);
                    float2 xa, xb, ua, ub;
                    r2c_split_pair(zA[r], zB[15 - r], w, xa, xb);
                    c2r_split_pair(cmul(xa, SA[r]), cmul(xb, SB[15 - r]), w, ua, ub);
                    qA[r] = conjf2(ua);
                    qB[15 - r] = conjf2(ub);
                }
            } else {
                const float2 pn = cmul(r2c_nyquist(zA[0]), Sn);
                qA[0] = conjf2(c2r_split(make_float2(0.f, 0.f), pn, make_float2(1.f, 0.f)));
                {
                    const float2 w8 = wkA(8);
                    const float2 p8 = cmul(r2c_split(zA[8], zA[8], w8), SA[8]);
                    qA[8] = conjf2(c2r_split(p8, p8, w8));
                }
#pragma unroll
                for (int r = 1; r < 8; ++r)
                    unit(zA[r], zA[16 - r], SA[r], SA[16 - r], wkA(r), wkA(16 - r), qA[r], qA[16 - r]);
#pragma unroll
                for (int r = 0; r < 8; ++r)
                    unit(zB[r], zB[15 - r], SB[r], SB[15 - r], wkB(r), wkB(15 - r), qB[r], qB[15 - r]);
            }
        }
        __syncthreads();  // all forward stage-3 reads are done before LDS is overwritten
        stage_first_to_lds<16>(qA, buf + m.sA * SEQS, m.jA);
        stage_first_to_lds<16>(qB, buf + m.sB * SEQS, m.jB);
        __syncthreads();
        stockham_stage<16, 16, ROW_LOG, 2, SEQS, NT, true>(buf, tti);
        {
            float2 F[32];
            stage_last_from_lds<32, N2>(F, buf + s1seq * SEQS, s1j, twl);
            // R'[row][t2] = F[t2] w_N^{t2 row}, t2 = j + 256 r:  w_L^{2 row j} * w_L^{512 row r}
            const uint32_t row = s1seq ? m.row1 : m.row0;
            const float2 base = tw_two_level(twlo, twhi, (2u * row * (uint32_t)s1j) & (g.L - 1u));
            const float2 step = tw_two_level(twlo, twhi, (512u * row) & (g.L - 1u));
            float2 anchor[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
                anchor[a] = tw_two_level(twlo, twhi, (512u * row * (uint32_t)(8 * (a + 1))) & (g.L - 1u));
            if (more) {
                mn = row_map(wn, N1, tid);
                wrAn = tw_two_level(twlo, twhi, mn.rowA);
                wrBn = tw_two_level(twlo, twhi, mn.rowB);
            }
            float2 tw = base;
            if (rt_tiled(g.lg1)) {
                // tiles: pair w, side s1seq, t2 = s1j + 256 r -> tile row t2 / RT_COLS = s1j / RT_COLS
                // + (256 / RT_COLS) r, one scalar offset per r
                const __amdgpu_buffer_rsrc_t ro = make_rsrc(Rt, N1 * N2 * 8u);
                const uint32_t vo = ((((uint32_t)s1j / RT_COLS) * (N1 >> 1) + w) * (2u * RT_COLS) +
                                     (uint32_t)s1seq * RT_COLS + ((uint32_t)s1j % RT_COLS)) * 8u;
                const uint32_t rstep = (256u / RT_COLS) * (N1 >> 1) * (2u * RT_COLS) * 8u;  // t2 -> t2 + 256
#pragma unroll
                for (int r = 0; r < 32; ++r) {
                    if ((r & 7) == 0 && r) tw = cmul(base, anchor[r / 8 - 1]);
                    buf_store_nt(cmul(F[r], tw), ro, vo, (uint32_t)r * rstep);
                    tw = cmul(tw, step);
                }
            } else {
                float2* op = Yk + (size_t)row * N2 + s1j;
#pragma unroll
                for (int r = 0; r < 32; ++r) {
                    if ((r & 7) == 0 && r)  // re-anchor every 8 steps to bound the recurrence error
                        tw = cmul(base, anchor[r / 8 - 1]);
                    st_nt(&op[256 * r], cmul(F[r], tw));
                    tw = cmul(tw, step);
                }
            }
        }
        __syncthreads();  // stage-3 reads done before the next item writes LDS
        if (!more) break;
        // next request: the first operand of the next item (issued here rather than during the
        // inverse transform, whose registers are full)
        s1_load(MODE == ROWS_STATIC ? Yk : Sst, wn, pre);
        w = wn;
        m = mn;
        wrA = wrAn;
        wrB = wrBn;
    }
}

// ------------------------------------------------------------------ pass A' ----
// Workgroup grp owns columns [W*grp, W*grp + W): G[t1] = FFT_N1(R'[.][t2])[t1] = conj of the
// unnormalised inverse, so c[2t] = Re G / N and c[2t+1] = -Im G / N at t = t2 + N2*t1.  R' is
// row-major, or pass B's tiles when rt_tiled(LG1).
// Writes one nibble (4 conv bits, MSB first) per (column pair, t1) inside the output window, in
// column-pair-major order nibT[pair * N1 + t1] so that each workgroup's writes are contiguous.
template <int LG1, bool BATCH>
__global__ __launch_bounds__(inv_cols_threads(LG1), rt_tiled(LG1) && LG1 <= RT_LG1 ? 2 : 1) void k_inv_cols(
    const float2* __restrict__ Rp, uint8_t* __restrict__ nibT, Geom g,
    const uint32_t* __restrict__ wk_cnt, const uint32_t* __restrict__ ws_cnt,
    const float2* __restrict__ tw14, const float2* __restrict__ twlo, const float2* __restrict__ twhi,
    float* __restrict__ margin, const BlockArgs* __restrict__ bt) {
    if constexpr (BATCH) {
        const BlockArgs& a = bt[blockIdx.z];
        g = a.g;
        Rp = rt_tiled(LG1) ? a.Rt : a.Yk;
        nibT = a.nib;
        wk_cnt = a.counts;
        ws_cnt = a.seed ? a.counts + POP_SLOTS : a.ws;
        margin = a.margin;
    }
    constexpr int N1 = 1 << LG1;
    constexpr int W = inv_cols_width(LG1);
    constexpr int NT = inv_cols_threads(LG1);
    constexpr int LGL = LG1 < 14 ? LG1 : 14;  // LDS transform length (N1 = 2^15: two halves)
    constexpr int SEQS = seq_stride(1 << LGL);
    constexpr int WL = (LG1 == RT_LG1 || LG1 >= 14) ? 1 : W;  // columns held in LDS at once
    __shared__ float2 buf[WL * SEQS];
    __shared__ float2 twl[TWL_SIZE];
    load_twl(twl, tw14, NT);
    const uint32_t grp = xcd_logical(blockIdx.x, gridDim.x);
    if constexpr (LG1 == 14) {
        // Column pair t2 = 2 grp, 2 grp + 1 per 1024-thread workgroup, one column in LDS at a time
        // (one workgroup per CU).  Thread j loads both columns of row k1 = j + 1024 r from pass B's
        // tiles with one 16-byte load (side 0 of pair k1 for k1 < N1/2, side 1 of pair N1 - k1
        // after).  The first column's bits wait packed in a register (2 per t1 = j + 1024 q) for the
        // second's; the output is one nibble per (column pair, t1).
        static_assert(W == 2 && NT == 1024 && rt_tiled(LG1), "paired column pass expects 2 columns, 1024 threads");
        constexpr int PT = N1 / NT;
        const int j = threadIdx.x;
        float2 v0[PT], v1[PT];
        const uint32_t c0 = grp * W;
        {
            const float2* tp = Rp + (size_t)(c0 / RT_COLS) * (N1 / 2) * (2 * RT_COLS) + (c0 % RT_COLS);
#pragma unroll
            for (int r = 0; r < PT; ++r) {
                const uint32_t k1 = (uint32_t)j + (uint32_t)NT * r;
                const uint32_t pr = r < PT / 2 ? k1 : (((uint32_t)N1 - k1) & ((uint32_t)N1 / 2 - 1));
                const float4 x = *reinterpret_cast<const float4*>(tp + (size_t)pr * (2 * RT_COLS) + (r < PT / 2 ? 0 : RT_COLS));
                v0[r] = make_float2(x.x, x.y);
                v1[r] = make_float2(x.z, x.w);
            }
        }
        const float corr = correction(pop_total(wk_cnt), pop_total(ws_cnt), g.L);
        const float invN = g.inv_N;
        float dist = 0.f;
        auto inwin = [&](uint32_t t1) {
            const uint32_t i0 = 2u * (c0 + g.N2 * t1);
            return !(i0 + 4u <= g.out_off || i0 >= g.out_off + g.out_bits);
        };
        __syncthreads();  // twl
#pragma unroll
        for (int r = 0; r < PT; ++r) buf[padi(j + NT * r)] = v0[r];
        __syncthreads();
        lds_fft<LG1, 1, SEQS, NT>(buf, twl);
        uint32_t lo = 0u;
#pragma unroll
        for (int q = 0; q < PT; ++q) {
            const uint32_t t1 = (uint32_t)j + (uint32_t)NT * q;
            if (!inwin(t1)) continue;
            const float2 G = buf[padi((int)t1)];
            const uint32_t b0 = round_bit(G.x * invN + corr, dist);
            const uint32_t b1 = round_bit(-G.y * invN + corr, dist);
            lo |= ((b0 << 1) | b1) << (2 * q);
        }
        __syncthreads();  // the first column's reads are done
#pragma unroll
        for (int r = 0; r < PT; ++r) buf[padi(j + NT * r)] = v1[r];
        __syncthreads();
        lds_fft<LG1, 1, SEQS, NT>(buf, twl);
#pragma unroll
        for (int q = 0; q < PT; ++q) {
            const uint32_t t1 = (uint32_t)j + (uint32_t)NT * q;
            if (!inwin(t1)) continue;
            const float2 G = buf[padi((int)t1)];
            const uint32_t b2 = round_bit(G.x * invN + corr, dist);
            const uint32_t b3 = round_bit(-G.y * invN + corr, dist);
            st_nt_u8(&nibT[(size_t)grp * N1 + t1], (uint8_t)((((lo >> (2 * q)) & 3u) << 2) | (b2 << 1) | b3));
        }
        if (margin) margin_publish(dist, margin);
        return;
    } else if constexpr (LG1 >= 14) {
        // One column t2 = grp per workgroup; output one 2-bit crumb per t1 at nibT[t2 * N1 + t1]
        // (bits c[2t], c[2t+1] of t = t2 + N2 t1).  N1 = 2^15 splits the column first by one
        // radix-2 step in registers (X[2k] = FFT(a)[k], X[2k+1] = FFT(b)[k] with a = x[j] + x[j+N1/2],
        // b = (x[j] - x[j+N1/2]) w_N1^j), then transforms the two 2^14-point halves in turn.
        static_assert(W == 1 && NT == 1024, "single-column pass expects 1024 threads");
        constexpr int H = 1 << LGL;       // points per LDS transform
        constexpr int PT = N1 / NT;       // values per thread (16 or 32)
        const uint32_t t2 = grp;
        const float corr = correction(pop_total(wk_cnt), pop_total(ws_cnt), g.L);
        const float invN = g.inv_N;
        float dist = 0.f;
        float2 v[PT];
        if constexpr (rt_tiled(LG1)) {
            // pass B's tiles: rows k1 < N1/2 are side 0 of pair k1, the others side 1 of pair
            // (N1 - k1) mod N1/2, so consecutive threads read consecutive tile lines
            const float2* tp = Rp + (size_t)(t2 / RT_COLS) * (N1 / 2) * (2 * RT_COLS) + (t2 % RT_COLS);
#pragma unroll
            for (int r = 0; r < PT; ++r) {
                const uint32_t k1 = threadIdx.x + (uint32_t)NT * r;
                const uint32_t pr = r < PT / 2 ? k1 : (((uint32_t)N1 - k1) & ((uint32_t)N1 / 2 - 1));
                v[r] = tp[(size_t)pr * (2 * RT_COLS) + (r < PT / 2 ? 0 : RT_COLS)];
            }
        } else {
#pragma unroll
            for (int r = 0; r < PT; ++r) v[r] = Rp[(size_t)(threadIdx.x + NT * r) * g.N2 + t2];
        }
        __syncthreads();  // twl
        constexpr int HALVES = N1 / H;
        if constexpr (HALVES == 2) {
            const uint32_t step = g.L / (uint32_t)N1;
#pragma unroll
            for (int r = 0; r < PT / 2; ++r) {
                const float2 x0 = v[r], x1 = v[r + PT / 2];
                v[r] = cadd(x0, x1);
                v[r + PT / 2] = cmul(csub(x0, x1), tw_two_level(twlo, twhi, (threadIdx.x + NT * r) * step));
            }
        }
#pragma unroll
        for (int hv = 0; hv < HALVES; ++hv) {
#pragma unroll
            for (int r = 0; r < H / NT; ++r) buf[padi(threadIdx.x + NT * r)] = v[hv * (H / NT) + r];
            __syncthreads();
            lds_fft<LGL, 1, SEQS, NT>(buf, twl);
            for (int k = threadIdx.x; k < H; k += NT) {
                const uint32_t t1 = (uint32_t)(k * HALVES + hv);
                const uint32_t i0 = 2u * (t2 + g.N2 * t1);
                if (i0 + 2u <= g.out_off || i0 >= g.out_off + g.out_bits) continue;
                const float2 G = buf[padi(k)];
                const uint32_t b0 = round_bit(G.x * invN + corr, dist);
                const uint32_t b1 = round_bit(-G.y * invN + corr, dist);
                st_nt_u8(&nibT[(size_t)t2 * N1 + t1], (uint8_t)((b0 << 1) | b1));
            }
            __syncthreads();  // reads done before the next half is written
        }
        if (margin) margin_publish(dist, margin);
        return;
    } else {
    if constexpr (LG1 == RT_LG1) {
        // One column pair t2 = 2 grp, 2 grp + 1 from pass B's tiles per 256-thread workgroup, one
        // column in LDS at a time (74 KB), so that two workgroups share a compute unit and one's
        // reads overlap the other's transforms.  Thread j loads both columns of row k1 = j + 256 r
        // with one 16-byte load: rows k1 < N1/2 are side 0 of pair k1, the others side 1 of pair
        // (N1 - k1) mod N1/2, so a workgroup walks consecutive lines.  Plan 32 x 16 x 16 per
        // column: stage 1 from the loaded registers, stage 2 in LDS, stage 3 into registers for
        // butterflies b = j and j + 256 (t1 = b + 512 r).  The first column's bits wait packed in
        // registers (2 per t1) for the second's.
        static_assert(W == 2 && NT == 256, "tiled column pass expects 2 columns, 256 threads");
        static_assert(RT_COLS % W == 0, "a workgroup's columns lie in one tile");
        __shared__ float2 ttf[32 * 17];
        fill_stage_table<16, 32>(ttf, tw14, NT);
        const int j = threadIdx.x;
        float2 v0[32], v1[32];
        {
            const uint32_t c0 = grp * W;
            const float2* tp = Rp + (size_t)(c0 / RT_COLS) * (N1 / 2) * (2 * RT_COLS) + (c0 % RT_COLS);
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                const uint32_t k1 = (uint32_t)j + 256u * r;
                const uint32_t pr = r < 16 ? k1 : ((N1 - k1) & (N1 / 2 - 1));
                const float4 x = *reinterpret_cast<const float4*>(tp + (size_t)pr * (2 * RT_COLS) + (r < 16 ? 0 : RT_COLS));
                v0[r] = make_float2(x.x, x.y);
                v1[r] = make_float2(x.z, x.w);
            }
        }
        const float corr = correction(pop_total(wk_cnt), pop_total(ws_cnt), g.L);
        const float invN = g.inv_N;
        float dist = 0.f;
        const uint32_t t2 = grp * W;
        auto inwin = [&](uint32_t t1) {
            const uint32_t i0 = 2u * (t2 + g.N2 * t1);
            return !(i0 + 4u <= g.out_off || i0 >= g.out_off + g.out_bits);
        };
        __syncthreads();  // twl, ttf
        stage_first_to_lds<32>(v0, buf, j);
        __syncthreads();
        stockham_stage<16, 32, LG1, 1, SEQS, NT, true>(buf, ttf);
        uint32_t lo[2] = {0u, 0u};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float2 G[16];
            const int bf = j + 256 * h;
            stage_last_from_lds<16, N1>(G, buf, bf, twl);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                if (!inwin((uint32_t)bf + 512u * r)) continue;
                const uint32_t b0 = round_bit(G[r].x * invN + corr, dist);
                const uint32_t b1 = round_bit(-G[r].y * invN + corr, dist);
                lo[h] |= ((b0 << 1) | b1) << (2 * r);
            }
        }
        __syncthreads();  // the first column's stage-3 reads are done
        stage_first_to_lds<32>(v1, buf, j);
        __syncthreads();
        stockham_stage<16, 32, LG1, 1, SEQS, NT, true>(buf, ttf);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float2 G[16];
            const int bf = j + 256 * h;
            stage_last_from_lds<16, N1>(G, buf, bf, twl);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const uint32_t t1 = (uint32_t)bf + 512u * r;
                if (!inwin(t1)) continue;
                const uint32_t b2 = round_bit(G[r].x * invN + corr, dist);
                const uint32_t b3 = round_bit(-G[r].y * invN + corr, dist);
                st_nt_u8(&nibT[(size_t)grp * N1 + t1], (uint8_t)((((lo[h] >> (2 * r)) & 3u) << 2) | (b2 << 1) | b3));
            }
        }
        if (margin) margin_publish(dist, margin);
        return;
    }
    constexpr int LPT = (W * N1 + NT - 1) / NT;
    float2 v[LPT];
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
        const int e = threadIdx.x + i * NT;
        if (e < W * N1) {
            const uint32_t k1 = (uint32_t)(e / W), t2 = grp * W + (uint32_t)(e % W);
            if constexpr (rt_tiled(LG1)) {  // pass B's tiles: side 0 of pair k1, or side 1 of N1 - k1
                const uint32_t pr = k1 < (uint32_t)N1 / 2 ? k1 : (((uint32_t)N1 - k1) & ((uint32_t)N1 / 2 - 1));
                const uint32_t sd = k1 < (uint32_t)N1 / 2 ? 0u : 1u;
                v[i] = Rp[((size_t)(t2 / RT_COLS) * (N1 / 2) + pr) * (2 * RT_COLS) + sd * RT_COLS + t2 % RT_COLS];
            } else {
                v[i] = Rp[(size_t)k1 * g.N2 + t2];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
        const int e = threadIdx.x + i * NT;
        if (e < W * N1) buf[(e % W) * SEQS + padi(e / W)] = v[i];
    }
    __syncthreads();
    lds_fft<LG1, W, SEQS, NT>(buf, twl);
    const float corr = correction(pop_total(wk_cnt), pop_total(ws_cnt), g.L);
    const float invN = g.inv_N;
    float dist = 0.f;
    constexpr int P = W / 2;
    // the output window [out_off, out_off + out_bits) as a t1 range for this pair of columns
    for (int e = threadIdx.x; e < P * N1; e += NT) {
        const int p = e / N1, t1 = e % N1;
        const uint32_t t2 = grp * W + 2u * (uint32_t)p;
        const uint32_t i0 = 2u * (t2 + g.N2 * (uint32_t)t1);
        if (i0 + 4u <= g.out_off || i0 >= g.out_off + g.out_bits) continue;
        const float2 G0 = buf[(2 * p) * SEQS + padi(t1)];
        const float2 G1 = buf[(2 * p + 1) * SEQS + padi(t1)];
        const uint32_t b0 = round_bit(G0.x * invN + corr, dist);
        const uint32_t b1 = round_bit(-G0.y * invN + corr, dist);
        const uint32_t b2 = round_bit(G1.x * invN + corr, dist);
        const uint32_t b3 = round_bit(-G1.y * invN + corr, dist);
        nibT[(size_t)(grp * P + (uint32_t)p) * N1 + (uint32_t)t1] = (uint8_t)((b0 << 3) | (b1 << 2) | (b2 << 1) | b3);
    }
    if (margin) margin_publish(dist, margin);
    }
}

// ------------------------------------------------------------------ output ----
// The conv bit stream in nibbles: nibble q (bits 4q..4q+3) is nibT[(q mod P2) * N1 + q / P2] with
// P2 = N2/2 column pairs.  A workgroup transposes a tile of PACK_T1 rows x PACK_G pairs through
// LDS and writes the output words whose 8 nibbles lie in the tile: conv bits
// [out_off + 32w, +32), XOR key_rest word w ((key[h/32+w] << 1) | (key[h/32+w+1] >> 31),
// key2StartRest PrivacyAmplification.cu:873-878), byte-swapped for the OutBytes layout
// (ToBinaryArray, PrivacyAmplification.cu:757-762).
// CRUMB: pass A' wrote one 2-bit crumb per column, crumb[t2 * N1 + t1]; the nibble of pair p is
// crumb[2p] << 2 | crumb[2p + 1].
// zero_counts (single-block form): the popcount partials, which no later stage of this block
// reads, are cleared for the next block on the stream.
template <bool BATCH, bool CRUMB>
__global__ __launch_bounds__(256) void k_pack(const uint8_t* __restrict__ nibT,
                                              const uint32_t* __restrict__ key,
                                              uint32_t* __restrict__ out, Geom g,
                                              const BlockArgs* __restrict__ bt,
                                              uint32_t* __restrict__ zero_counts) {
    if (!BATCH && zero_counts && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 2 * POP_SLOTS)
        zero_counts[threadIdx.x] = 0u;
    __shared__ uint8_t tile[PACK_T1][PACK_G + 4];
    if constexpr (BATCH) {  // grid.y covers the largest window of the launch; rows past this
        const BlockArgs& a = bt[blockIdx.z];  // block's window load nothing and store nothing
        g = a.g;
        nibT = a.nib;
        key = a.key;
        out = a.out;
    }
    const uint32_t P2 = g.N2 >> 1;
    const uint32_t q_lo = g.out_off >> 2, q_hi = (g.out_off + g.out_bits) >> 2;
    const uint32_t row0 = ((q_lo / P2) & ~3u) + blockIdx.y * PACK_T1;  // 4-aligned for 32-bit reads
    const uint32_t g0 = blockIdx.x * PACK_G;
    // load: PACK_G pairs x PACK_T1 rows, 4 rows per 32-bit read
    for (uint32_t e = threadIdx.x; e < PACK_G * (PACK_T1 / 4); e += 256) {
        const uint32_t gp = e / (PACK_T1 / 4), r4 = (e % (PACK_T1 / 4)) * 4;
        const uint32_t row = row0 + r4;
        uint32_t v = 0;
        if constexpr (CRUMB) {
            if (row < g.N1) {
                const uint32_t ve = *reinterpret_cast<const uint32_t*>(nibT + (size_t)(2u * (g0 + gp)) * g.N1 + row);
                const uint32_t vo = *reinterpret_cast<const uint32_t*>(nibT + (size_t)(2u * (g0 + gp) + 1u) * g.N1 + row);
                v = ((ve & 0x03030303u) << 2) | (vo & 0x03030303u);
            }
        } else {
            if (row < g.N1) v = *reinterpret_cast<const uint32_t*>(nibT + (size_t)(g0 + gp) * g.N1 + row);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) tile[r4 + k][gp] = (uint8_t)(v >> (8 * k));
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < PACK_T1 * (PACK_G / 8); e += 256) {
        const uint32_t r = e / (PACK_G / 8), gw = (e % (PACK_G / 8)) * 8;
        const uint32_t q = (row0 + r) * P2 + g0 + gw;
        if (q < q_lo || q >= q_hi) continue;
        const uint32_t w = (q - q_lo) >> 3;
        uint32_t cw = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) cw |= ((uint32_t)tile[r][gw + j] & 0xFu) << (28 - 4 * j);
        if (g.do_xor && 32u * w < g.m) cw ^= (key[g.key_hw + w] << 1) | (key[g.key_hw + w + 1] >> 31);
        out[w] = g.reverse_endian ? __builtin_bswap32(cw) : cw;
    }
}

// ------------------------------------------------------------------ small path ----
// The whole block in one workgroup: N = 2^LGN <= 8192 complex points per operand in LDS.
// sspec != nullptr: static seed (precomputed split spectrum, N+1 values, popcount *ws_ptr).
// SEEDPREP: only computes the seed's split spectrum into sspec_out and its popcount.
template <int LGN, bool SEEDPREP, bool BATCH>
__global__ __launch_bounds__(small_threads(LGN)) void k_small(
    const uint32_t* __restrict__ key, const uint32_t* __restrict__ seed,
    const float2* __restrict__ sspec, const uint32_t* __restrict__ ws_ptr,
    uint32_t* __restrict__ out, Geom g, const float2* __restrict__ tw14,
    float* __restrict__ margin, float2* __restrict__ sspec_out, uint32_t* __restrict__ ws_out,
    const BlockArgs* __restrict__ bt) {
    if constexpr (BATCH) {
        const BlockArgs& a = bt[blockIdx.z];
        g = a.g;
        key = a.key;
        seed = a.seed;
        sspec = a.seed ? nullptr : a.sspec;
        ws_ptr = a.ws;
        out = a.out;
        margin = a.margin;
    }
    constexpr int N = 1 << LGN, L = 2 * N;
    constexpr int NT = small_threads(LGN);
    constexpr int SEQS = seq_stride(N);
    constexpr int PER = (N + NT - 1) / NT;
    __shared__ float2 buf[2 * SEQS];
    __shared__ float2 twl[TWL_SIZE];
    __shared__ uint32_t cnt[2];
    const int tid = threadIdx.x;
    load_twl(twl, tw14, NT);
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
        lds_fft<LGN, 2, SEQS, NT>(buf, twl);  // sequence 0 is unused
    } else if (dyn) {
        lds_fft<LGN, 2, SEQS, NT>(buf, twl);
    } else {
        lds_fft<LGN, 1, SEQS, NT>(buf, twl);
    }
    constexpr int TWS = TW14 / L;
    float2 X[PER];
    float2 PN = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int k = tid + i * NT;
        if (k >= N) continue;
        const int kp = (N - k) & (N - 1);
        const float2 wk = tw_lds(twl, k * TWS);
        float2 s, sn = make_float2(0.f, 0.f);
        if (dyn) {
            const float2 zk = buf[SEQS + padi(k)], zp = buf[SEQS + padi(kp)];
            s = r2c_split(zk, zp, wk);
            if (k == 0) {
                sn = r2c_nyquist(zk);
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
            PN = cmul(r2c_nyquist(zk), sn);
            x = make_float2(0.f, 0.f);
        }
        X[i] = cmul(x, s);
    }
    if (SEEDPREP) {
        if (tid < POP_SLOTS) ws_out[tid] = (tid == 0) ? cnt[1] : 0u;
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
        X[i] = conjf2(c2r_split(X[i], pp, tw_lds(twl, k * TWS)));
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int k = tid + i * NT;
        if (k < N) buf[padi(k)] = X[i];
    }
    __syncthreads();
    lds_fft<LGN, 1, SEQS, NT>(buf, twl);
    const uint32_t ws = dyn ? cnt[1] : pop_total(ws_ptr);
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

// ------------------------------------------------------------------ sub-block split ----
// The m x (h+1) Toeplitz product cut into b x b tiles (the tiling of LargeBlocksizeExample,
// LBS/main.cpp:104-124, 313-523): tile (band i, key column j) is the length-2b circular product of
// key bits [jb, jb+b) (bit b zero) with the seed segment S[R, R+b) || S[R-b, R) at R = (i-j)b mod n,
// read at [0, b): z[v] = XOR_u x[jb+u] s[((i-j)b + v - u) mod n].  Tiles of a chunk are consecutive
// key columns j0.., all bands each: tile t = (j - j0) * nbands + i.
__global__ __launch_bounds__(256) void k_tile_inputs(const uint32_t* __restrict__ key,
                                                     const uint32_t* __restrict__ seed, uint32_t n,
                                                     uint32_t h, uint32_t b, uint32_t nbands, uint32_t j0,
                                                     uint32_t* __restrict__ tkeys, uint32_t* __restrict__ tseeds) {
    const uint32_t t = blockIdx.y, i = t % nbands, j = j0 + t / nbands;
    const uint32_t wb = b >> 5, nw = n >> 5, hw = h >> 5;
    const uint32_t R = (uint32_t)(((int64_t)i - (int64_t)j) * (int64_t)b % (int64_t)n + (int64_t)n) % n;
    const uint32_t R2 = (R + n - b % n) % n;
    uint32_t* tk = tkeys + (size_t)t * (wb + 1);
    uint32_t* ts = tseeds + (size_t)t * (2 * wb);
    for (uint32_t w = blockIdx.x * 256u + threadIdx.x; w < 3u * wb + 1u; w += gridDim.x * 256u) {
        if (w <= wb) {  // key words; word wb (bit b) is zero
            const uint32_t kw = j * wb + w;
            uint32_t v = 0;
            if (w < wb) v = kw < hw ? key[kw] : (kw == hw ? key[kw] & 0x80000000u : 0u);
            tk[w] = v;
        } else {
            const uint32_t sw = w - wb - 1u;  // seed word of the tile, < 2 wb
            const uint32_t src = sw < wb ? (R >> 5) + sw : (R2 >> 5) + (sw - wb);
            ts[sw] = seed[src % nw];
        }
    }
}

// acc[w] (=|^=) XOR over the chunk's columns of the tile outputs of band w / wb (32-bit words of
// the engine's output layout, which XOR word for word).
__global__ __launch_bounds__(256) void k_tile_xor(uint32_t* __restrict__ acc, const uint32_t* __restrict__ tout,
                                                  uint32_t nbands, uint32_t wb, uint32_t cols, int first) {
    const uint32_t w = blockIdx.x * 256u + threadIdx.x;
    if (w >= nbands * wb) return;
    const uint32_t i = w / wb, ww = w % wb;
    uint32_t x = first ? 0u : acc[w];
    for (uint32_t c = 0; c < cols; ++c) x ^= tout[(size_t)(c * nbands + i) * wb + ww];
    acc[w] = x;
}

// out word w = acc word w XOR the key rest word ((key[h/32+w] << 1) | (key[h/32+w+1] >> 31),
// key2StartRest PrivacyAmplification.cu:873-878) in the output byte order.
__global__ __launch_bounds__(256) void k_split_finish(const uint32_t* __restrict__ acc,
                                                      const uint32_t* __restrict__ key, uint32_t* __restrict__ out,
                                                      uint32_t m, uint32_t h, uint32_t do_xor, uint32_t reverse_endian) {
    const uint32_t w = blockIdx.x * 256u + threadIdx.x;
    if (w >= (m >> 5)) return;
    uint32_t kr = 0;
    if (do_xor) kr = (key[(h >> 5) + w] << 1) | (key[(h >> 5) + w + 1] >> 31);
    out[w] = acc[w] ^ (reverse_endian ? __builtin_bswap32(kr) : kr);
}

hipError_t launch_tile_inputs(const uint32_t* key, const uint32_t* seed, uint32_t n, uint32_t h, uint32_t b,
                              uint32_t nbands, uint32_t j0, uint32_t tiles, uint32_t* tkeys, uint32_t* tseeds,
                              hipStream_t s) {
    const uint32_t words = 3u * (b >> 5) + 1u;
    k_tile_inputs<<<dim3((words + 255u) / 256u, tiles), 256, 0, s>>>(key, seed, n, h, b, nbands, j0, tkeys, tseeds);
    return hipGetLastError();
}
hipError_t launch_tile_xor(uint32_t* acc, const uint32_t* tout, uint32_t nbands, uint32_t wb, uint32_t cols,
                           int first, hipStream_t s) {
    k_tile_xor<<<(nbands * wb + 255u) / 256u, 256, 0, s>>>(acc, tout, nbands, wb, cols, first);
    return hipGetLastError();
}
hipError_t launch_split_finish(const uint32_t* acc, const uint32_t* key, uint32_t* out, uint32_t m, uint32_t h,
                               uint32_t do_xor, uint32_t reverse_endian, hipStream_t s) {
    k_split_finish<<<((m >> 5) + 255u) / 256u, 256, 0, s>>>(acc, key, out, m, h, do_xor, reverse_endian);
    return hipGetLastError();
}

// ================================================================== launchers ====
namespace {
template <int LGM, int D>
hipError_t launch_fwd_cols_t(const uint32_t* tb, float2* Y, const Geom& g, const Tables& t, hipStream_t s,
                             const uint32_t* tb2, float2* Y2) {
    if constexpr (LGM == 9) {
        // Each workgroup holds one class of one operand and walks `per` column groups: 16 (8 for a
        // single operand), fewer while that would leave under 256 (512) workgroups per operand.
        // 1024 workgroups at 2^27, two waves of two per CU: as fast as one persistent wave, and a
        // CU that cannot take its second workgroup at once delays a short walk, not a long one.
        const uint32_t groups = g.N2 / 16u, target = Y2 ? 256u : 512u;
        uint32_t per = Y2 ? 16u : 8u;
        while (per > 1u && groups * (uint32_t)D / per < target) per >>= 1;
        const uint32_t G = groups * (uint32_t)D / per;  // groups = 512 and per divides it
        k_fwd_cols9p<D, false><<<dim3(G, Y2 ? 2 : 1), FWD_NT, 0, s>>>(tb, Y, g, t.tw14, t.lo, t.hi, t.cls, tb2, Y2,
                                                                     per, nullptr, 0);
    } else {
        const uint32_t grid = (g.N2 / (uint32_t)fwd_cols_width(LGM)) * (uint32_t)D;
        k_fwd_cols<LGM, D, false><<<dim3(grid, Y2 ? 2 : 1), fwd_cols_threads(LGM), 0, s>>>(tb, Y, g, t.tw14, t.lo, t.hi, t.cls,
                                                                        nullptr, 0, tb2, Y2);
    }
    return hipGetLastError();
}
template <int LG1>
hipError_t launch_inv_cols_t(const float2* R, uint8_t* nib, const Geom& g, const uint32_t* wk,
                             const uint32_t* ws, const Tables& t, float* margin, hipStream_t s) {
    const uint32_t grid = g.N2 / (uint32_t)inv_cols_width(LG1);
    k_inv_cols<LG1, false><<<grid, inv_cols_threads(LG1), 0, s>>>(R, nib, g, wk, ws, t.tw14, t.lo, t.hi, margin, nullptr);
    return hipGetLastError();
}
template <int D>
hipError_t launch_transpose_t(const uint32_t* bits, uint32_t* tb, uint32_t* count, const Geom& g,
                              int is_seed, hipStream_t s, const uint32_t* bits2, uint32_t* tb2) {
    const dim3 grid(g.M << (ROW_LOG - TB_SPAN_LOG), bits2 ? 2 : 1);
    k_transpose_bits<D, false><<<grid, TB_NT, 0, s>>>(bits, tb, count, g, is_seed, nullptr, bits2, tb2);
    return hipGetLastError();
}
template <int LGN>
hipError_t launch_small_t(const uint32_t* key, const uint32_t* seed, const float2* sspec,
                          const uint32_t* ws_ptr, uint32_t* out, const Geom& g, const Tables& t,
                          float* margin, float2* sspec_out, uint32_t* ws_out, hipStream_t s) {
    if (sspec_out)
        k_small<LGN, true, false><<<1, small_threads(LGN), 0, s>>>(key, seed, sspec, ws_ptr, out, g, t.tw14,
                                                                  margin, sspec_out, ws_out, nullptr);
    else
        k_small<LGN, false, false><<<1, small_threads(LGN), 0, s>>>(key, seed, sspec, ws_ptr, out, g, t.tw14,
                                                                   margin, sspec_out, ws_out, nullptr);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_copies(const CopyDesc* d, uint32_t count, uint64_t max_bytes, hipStream_t s) {
    if (count == 0) return hipSuccess;
    const uint64_t chunks = (max_bytes + COPY_CHUNK - 1) / COPY_CHUNK;
    const uint32_t y = (uint32_t)(chunks < 1 ? 1 : (chunks > 32 ? 32 : chunks));
    k_copies<<<dim3(count, y), COPY_NT, 0, s>>>(d);
    return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t* p, uint32_t count, hipStream_t s) {
    k_zero_u32<false><<<1, 256, 0, s>>>(p, count, nullptr);
    return hipGetLastError();
}

hipError_t launch_transpose(const uint32_t* bits, uint32_t* tb, uint32_t* count, const Geom& g,
                            int is_seed, hipStream_t s, const uint32_t* bits2, uint32_t* tb2) {
    switch (g.D) {
    case 1: return launch_transpose_t<1>(bits, tb, count, g, is_seed, s, bits2, tb2);
    case 2: return launch_transpose_t<2>(bits, tb, count, g, is_seed, s, bits2, tb2);
    case 4: return launch_transpose_t<4>(bits, tb, count, g, is_seed, s, bits2, tb2);
    case 8: return launch_transpose_t<8>(bits, tb, count, g, is_seed, s, bits2, tb2);
    case 16: return launch_transpose_t<16>(bits, tb, count, g, is_seed, s, bits2, tb2);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fwd_cols(const uint32_t* tb, float2* Y, const Geom& g, const Tables& t, hipStream_t s,
                           const uint32_t* tb2, float2* Y2) {
    switch (g.lg1) {
    case 1: return launch_fwd_cols_t<1, 1>(tb, Y, g, t, s, tb2, Y2);
    case 2: return launch_fwd_cols_t<2, 1>(tb, Y, g, t, s, tb2, Y2);
    case 3: return launch_fwd_cols_t<3, 1>(tb, Y, g, t, s, tb2, Y2);
    case 4: return launch_fwd_cols_t<4, 1>(tb, Y, g, t, s, tb2, Y2);
    case 5: return launch_fwd_cols_t<5, 1>(tb, Y, g, t, s, tb2, Y2);
    case 6: return launch_fwd_cols_t<6, 1>(tb, Y, g, t, s, tb2, Y2);
    case 7: return launch_fwd_cols_t<7, 1>(tb, Y, g, t, s, tb2, Y2);
    case 8: return launch_fwd_cols_t<8, 1>(tb, Y, g, t, s, tb2, Y2);
    case 9: return launch_fwd_cols_t<9, 1>(tb, Y, g, t, s, tb2, Y2);
    case 10: return launch_fwd_cols_t<9, 2>(tb, Y, g, t, s, tb2, Y2);
    case 11: return launch_fwd_cols_t<9, 4>(tb, Y, g, t, s, tb2, Y2);
    case 12: return launch_fwd_cols_t<9, 8>(tb, Y, g, t, s, tb2, Y2);
    case 13: return launch_fwd_cols_t<9, 16>(tb, Y, g, t, s, tb2, Y2);
    case 14: return launch_fwd_cols_t<10, 16>(tb, Y, g, t, s, tb2, Y2);
    case 15: return launch_fwd_cols_t<11, 16>(tb, Y, g, t, s, tb2, Y2);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rows(int mode, float2* Yk, float2* S, float2* Rt, const Geom& g, const Tables& t, int cus,
                       hipStream_t s) {
    const uint32_t items = g.N1 / 2u;
    const uint32_t grid = items < (uint32_t)cus ? items : (uint32_t)cus;  // one workgroup per CU
    switch (mode) {
    case ROWS_STATIC: k_rows<ROWS_STATIC, false><<<grid, ROWS_NT, 0, s>>>(Yk, S, Rt, g, t.tw14, t.lo, t.hi, nullptr); break;
    case ROWS_DYN: k_rows<ROWS_DYN, false><<<grid, ROWS_NT, 0, s>>>(Yk, S, Rt, g, t.tw14, t.lo, t.hi, nullptr); break;
    case ROWS_SEEDPREP: k_rows<ROWS_SEEDPREP, false><<<grid, ROWS_NT, 0, s>>>(Yk, S, Rt, g, t.tw14, t.lo, t.hi, nullptr); break;
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
    case 14: return launch_inv_cols_t<14>(R, nib, g, wk, ws, t, margin, s);
    case 15: return launch_inv_cols_t<15>(R, nib, g, wk, ws, t, margin, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_pack(const uint8_t* nibT, const uint32_t* key, uint32_t* out, const Geom& g, hipStream_t s,
                       uint32_t* zero_counts) {
    const uint32_t P2 = g.N2 >> 1;
    const uint32_t q_lo = g.out_off >> 2, q_hi = (g.out_off + g.out_bits) >> 2;
    const uint32_t rows = (q_hi - 1u) / P2 - ((q_lo / P2) & ~3u) + 1u;
    dim3 grid(P2 / PACK_G, (rows + PACK_T1 - 1u) / PACK_T1);
    if (g.W == 1) k_pack<false, true><<<grid, 256, 0, s>>>(nibT, key, out, g, nullptr, zero_counts);
    else k_pack<false, false><<<grid, 256, 0, s>>>(nibT, key, out, g, nullptr, zero_counts);
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

// ---------------------------------------------------------------- batched launchers ----
// `count` blocks of one transform length; block b's geometry and buffers are bt[b] (device
// memory) and g0 is any block's geometry (only its L-derived fields are used here).
uint32_t pack_tiles(const Geom& g) {
    const uint32_t P2 = g.N2 >> 1;
    const uint32_t q_lo = g.out_off >> 2, q_hi = (g.out_off + g.out_bits) >> 2;
    const uint32_t rows = (q_hi - 1u) / P2 - ((q_lo / P2) & ~3u) + 1u;
    return (rows + PACK_T1 - 1u) / PACK_T1;
}

hipError_t launch_batch_zero(const BlockArgs* bt, uint32_t count, hipStream_t s) {
    k_zero_u32<true><<<dim3(1, 1, count), 256, 0, s>>>(nullptr, 2 * POP_SLOTS, bt);
    return hipGetLastError();
}

namespace {
template <int D>
hipError_t batch_transpose_t(const BlockArgs* bt, uint32_t count, const Geom& g0, int is_seed, hipStream_t s) {
    k_transpose_bits<D, true><<<dim3(g0.M << (ROW_LOG - TB_SPAN_LOG), 1, count), TB_NT, 0, s>>>(
        nullptr, nullptr, nullptr, g0, is_seed, bt, nullptr, nullptr);
    return hipGetLastError();
}
template <int LGM, int D>
hipError_t batch_fwd_cols_t(const BlockArgs* bt, uint32_t count, const Geom& g0, const Tables& t, int is_seed,
                            uint32_t per, hipStream_t s) {
    if constexpr (LGM == 9) {
        // one class per workgroup walking `per` column groups (a power of two dividing the N2 / 16
        // groups); by default 16, fewer while the launch would have under 512 workgroups
        const uint32_t groups = g0.N2 / 16u;
        if (per == 0u) {
            per = 16u;
            while (per > 1u && groups * (uint32_t)D / per * count < 512u) per >>= 1;
        }
        if (per > FWD_PER_MAX || (per & (per - 1u)) != 0u || groups % per != 0u) return hipErrorInvalidValue;
        const uint32_t G = groups * (uint32_t)D / per;
        k_fwd_cols9p<D, true><<<dim3(G, 1, count), FWD_NT, 0, s>>>(nullptr, nullptr, g0, t.tw14, t.lo, t.hi, t.cls,
                                                                  nullptr, nullptr, per, bt, is_seed);
    } else {
        k_fwd_cols<LGM, D, true><<<dim3((g0.N2 / (uint32_t)fwd_cols_width(LGM)) * (uint32_t)D, 1, count),
                                   fwd_cols_threads(LGM), 0,
                                   s>>>(nullptr, nullptr, g0, t.tw14, t.lo, t.hi, t.cls, bt, is_seed, nullptr, nullptr);
    }
    return hipGetLastError();
}
template <int LG1>
hipError_t batch_inv_cols_t(const BlockArgs* bt, uint32_t count, const Geom& g0, const Tables& t, hipStream_t s) {
    k_inv_cols<LG1, true><<<dim3(g0.N2 / (uint32_t)inv_cols_width(LG1), 1, count), inv_cols_threads(LG1), 0, s>>>(
        nullptr, nullptr, g0, nullptr, nullptr, t.tw14, t.lo, t.hi, nullptr, bt);
    return hipGetLastError();
}
template <int LGN>
hipError_t batch_small_t(const BlockArgs* bt, uint32_t count, const Geom& g0, const Tables& t, hipStream_t s) {
    k_small<LGN, false, true><<<dim3(1, 1, count), small_threads(LGN), 0, s>>>(
        nullptr, nullptr, nullptr, nullptr, nullptr, g0, t.tw14, nullptr, nullptr, nullptr, bt);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_batch_transpose(const BlockArgs* bt, uint32_t count, const Geom& g0, int is_seed, hipStream_t s) {
    switch (g0.D) {
    case 1: return batch_transpose_t<1>(bt, count, g0, is_seed, s);
    case 2: return batch_transpose_t<2>(bt, count, g0, is_seed, s);
    case 4: return batch_transpose_t<4>(bt, count, g0, is_seed, s);
    case 8: return batch_transpose_t<8>(bt, count, g0, is_seed, s);
    case 16: return batch_transpose_t<16>(bt, count, g0, is_seed, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch_fwd_cols(const BlockArgs* bt, uint32_t count, const Geom& g0, const Tables& t, int is_seed,
                                 uint32_t per, hipStream_t s) {
    switch (g0.lg1) {
    case 1: return batch_fwd_cols_t<1, 1>(bt, count, g0, t, is_seed, per, s);
    case 2: return batch_fwd_cols_t<2, 1>(bt, count, g0, t, is_seed, per, s);
    case 3: return batch_fwd_cols_t<3, 1>(bt, count, g0, t, is_seed, per, s);
    case 4: return batch_fwd_cols_t<4, 1>(bt, count, g0, t, is_seed, per, s);
    case 5: return batch_fwd_cols_t<5, 1>(bt, count, g0, t, is_seed, per, s);
    case 6: return batch_fwd_cols_t<6, 1>(bt, count, g0, t, is_seed, per, s);
    case 7: return batch_fwd_cols_t<7, 1>(bt, count, g0, t, is_seed, per, s);
    case 8: return batch_fwd_cols_t<8, 1>(bt, count, g0, t, is_seed, per, s);
    case 9: return batch_fwd_cols_t<9, 1>(bt, count, g0, t, is_seed, per, s);
    case 10: return batch_fwd_cols_t<9, 2>(bt, count, g0, t, is_seed, per, s);
    case 11: return batch_fwd_cols_t<9, 4>(bt, count, g0, t, is_seed, per, s);
    case 12: return batch_fwd_cols_t<9, 8>(bt, count, g0, t, is_seed, per, s);
    case 13: return batch_fwd_cols_t<9, 16>(bt, count, g0, t, is_seed, per, s);
    case 14: return batch_fwd_cols_t<10, 16>(bt, count, g0, t, is_seed, per, s);
    case 15: return batch_fwd_cols_t<11, 16>(bt, count, g0, t, is_seed, per, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch_rows(int mode, const BlockArgs* bt, uint32_t count, const Geom& g0, const Tables& t,
                             uint32_t grid_x, hipStream_t s) {
    const uint32_t items = g0.N1 / 2u;
    if (grid_x == 0 || grid_x > items) grid_x = items;
    const dim3 grid(grid_x, 1, count);
    switch (mode) {
    case ROWS_STATIC: k_rows<ROWS_STATIC, true><<<grid, ROWS_NT, 0, s>>>(nullptr, nullptr, nullptr, g0, t.tw14, t.lo, t.hi, bt); break;
    case ROWS_DYN: k_rows<ROWS_DYN, true><<<grid, ROWS_NT, 0, s>>>(nullptr, nullptr, nullptr, g0, t.tw14, t.lo, t.hi, bt); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_batch_inv_cols(const BlockArgs* bt, uint32_t count, const Geom& g0, const Tables& t, hipStream_t s) {
    switch (g0.lg1) {
    case 1: return batch_inv_cols_t<1>(bt, count, g0, t, s);
    case 2: return batch_inv_cols_t<2>(bt, count, g0, t, s);
    case 3: return batch_inv_cols_t<3>(bt, count, g0, t, s);
    case 4: return batch_inv_cols_t<4>(bt, count, g0, t, s);
    case 5: return batch_inv_cols_t<5>(bt, count, g0, t, s);
    case 6: return batch_inv_cols_t<6>(bt, count, g0, t, s);
    case 7: return batch_inv_cols_t<7>(bt, count, g0, t, s);
    case 8: return batch_inv_cols_t<8>(bt, count, g0, t, s);
    case 9: return batch_inv_cols_t<9>(bt, count, g0, t, s);
    case 10: return batch_inv_cols_t<10>(bt, count, g0, t, s);
    case 11: return batch_inv_cols_t<11>(bt, count, g0, t, s);
    case 12: return batch_inv_cols_t<12>(bt, count, g0, t, s);
    case 13: return batch_inv_cols_t<13>(bt, count, g0, t, s);
    case 14: return batch_inv_cols_t<14>(bt, count, g0, t, s);
    case 15: return batch_inv_cols_t<15>(bt, count, g0, t, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch_pack(const BlockArgs* bt, uint32_t count, const Geom& g0, uint32_t tiles_y, hipStream_t s) {
    const dim3 grid((g0.N2 >> 1) / PACK_G, tiles_y, count);
    if (g0.W == 1) k_pack<true, true><<<grid, 256, 0, s>>>(nullptr, nullptr, nullptr, g0, bt, nullptr);
    else k_pack<true, false><<<grid, 256, 0, s>>>(nullptr, nullptr, nullptr, g0, bt, nullptr);
    return hipGetLastError();
}

hipError_t launch_batch_small(const BlockArgs* bt, uint32_t count, const Geom& g0, const Tables& t, hipStream_t s) {
    switch (g0.lgN) {
    case 5: return batch_small_t<5>(bt, count, g0, t, s);
    case 6: return batch_small_t<6>(bt, count, g0, t, s);
    case 7: return batch_small_t<7>(bt, count, g0, t, s);
    case 8: return batch_small_t<8>(bt, count, g0, t, s);
    case 9: return batch_small_t<9>(bt, count, g0, t, s);
    case 10: return batch_small_t<10>(bt, count, g0, t, s);
    case 11: return batch_small_t<11>(bt, count, g0, t, s);
    case 12: return batch_small_t<12>(bt, count, g0, t, s);
    case 13: return batch_small_t<13>(bt, count, g0, t, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pa

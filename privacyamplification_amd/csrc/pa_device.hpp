// pa_device.hpp -- device building blocks of the gfx950 Toeplitz-hashing engine.
//
// Complex arithmetic, in-register DFTs of size 2..32, and an in-place Stockham FFT over LDS.
// All transforms are forward (w = exp(-2 pi i / L)) and unnormalized; inverses are taken as
// conj(FFT(conj(x))) by the callers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pa {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * (-i)
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

// exp(-2 pi i k / 32), k in [0, 16): exact constants rounded to float.
__host__ __device__ constexpr float w32_re(int k) {
    return k == 0 ? 1.0f : k == 1 ? 9.807852804e-01f : k == 2 ? 9.238795325e-01f
         : k == 3 ? 8.314696123e-01f : k == 4 ? 7.071067812e-01f : k == 5 ? 5.555702330e-01f
         : k == 6 ? 3.826834324e-01f : k == 7 ? 1.950903220e-01f : k == 8 ? 0.0f
         : k == 9 ? -1.950903220e-01f : k == 10 ? -3.826834324e-01f : k == 11 ? -5.555702330e-01f
         : k == 12 ? -7.071067812e-01f : k == 13 ? -8.314696123e-01f : k == 14 ? -9.238795325e-01f
         : -9.807852804e-01f;
}
__host__ __device__ constexpr float w32_im(int k) {
    return k == 0 ? 0.0f : k == 1 ? -1.950903220e-01f : k == 2 ? -3.826834324e-01f
         : k == 3 ? -5.555702330e-01f : k == 4 ? -7.071067812e-01f : k == 5 ? -8.314696123e-01f
         : k == 6 ? -9.238795325e-01f : k == 7 ? -9.807852804e-01f : k == 8 ? -1.0f
         : k == 9 ? -9.807852804e-01f : k == 10 ? -9.238795325e-01f : k == 11 ? -8.314696123e-01f
         : k == 12 ? -7.071067812e-01f : k == 13 ? -5.555702330e-01f : k == 14 ? -3.826834324e-01f
         : -1.950903220e-01f;
}

// v * exp(-2 pi i k / R) for a compile-time-foldable k in [0, R/2), R <= 32.
template <int R>
__device__ __forceinline__ float2 twc(float2 v, int k) {
    const int k32 = k * (32 / R);
    if (k32 == 0) return v;
    if (k32 == 8) return mul_mi(v);
    return cmul(v, make_float2(w32_re(k32), w32_im(k32)));
}

// v * exp(-2 pi i r / 32) for a compile-time-foldable r in [0, 32).
__device__ __forceinline__ float2 w32pow(float2 v, int r) {
    if (r < 16) return twc<32>(v, r);
    const float2 t = twc<32>(v, r - 16);
    return make_float2(-t.x, -t.y);
}

// tan(pi q / 16), q = 1, 2, 3, 4 (exact values rounded to float).
__host__ __device__ constexpr float tan16(int q) {
    return q == 1 ? 1.989123674e-01f : q == 2 ? 4.142135624e-01f : q == 3 ? 6.681786379e-01f : 1.0f;
}
// For w = exp(-2 pi i k / 32) = c + i s, k in [1, 16): the ratio of magnitude <= 1, s / c where
// |c| >= |s| (k <= 4 or k >= 12), else c / s.
__host__ __device__ constexpr bool w32_cform(int k) { return k <= 4 || k >= 12; }
__host__ __device__ constexpr float w32_ratio(int k) {
    return k <= 4 ? -tan16(k) : k < 8 ? -tan16(8 - k) : k < 12 ? tan16(k - 8) : tan16(16 - k);
}

// The butterfly (e + w o, e - w o), w = exp(-2 pi i k / 32) for a compile-time-foldable k in
// [1, 16) other than 8, in six FMAs instead of a complex product and four additions:
// w o = c (o + i t o) with t = s / c, or s (u o + i o) with u = c / s, whichever ratio is at
// most 1 in magnitude (the tangent form of the twiddled butterfly).
__device__ __forceinline__ void bfly_w32(float2 e, float2 o, int k, float2& lo, float2& hi) {
    const float r = w32_ratio(k);
    float2 z;
    float m;
    if (w32_cform(k)) {
        z = make_float2(fmaf(-r, o.y, o.x), fmaf(r, o.x, o.y));
        m = w32_re(k);
    } else {
        z = make_float2(fmaf(r, o.x, -o.y), fmaf(r, o.y, o.x));
        m = w32_im(k);
    }
    lo = make_float2(fmaf(m, z.x, e.x), fmaf(m, z.y, e.y));
    hi = make_float2(fmaf(-m, z.x, e.x), fmaf(-m, z.y, e.y));
}

// In-register forward DFT, natural-order input and output (radix-2 decimation in time).
template <int R>
__device__ __forceinline__ void dft(float2 (&x)[R]) {
    if constexpr (R == 2) {
        const float2 a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const float2 t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
        const float2 t2 = cadd(x[1], x[3]), t3 = mul_mi(csub(x[1], x[3]));
        x[0] = cadd(t0, t2);
        x[2] = csub(t0, t2);
        x[1] = cadd(t1, t3);
        x[3] = csub(t1, t3);
    } else if constexpr (R > 4) {
        constexpr int Q = R / 2;
        float2 e[Q], o[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            e[i] = x[2 * i];
            o[i] = x[2 * i + 1];
        }
        dft<Q>(e);
        dft<Q>(o);
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            const int k32 = k * (32 / R);
            if (k32 == 0 || k32 == 8) {
                const float2 t = twc<R>(o[k], k);
                x[k] = cadd(e[k], t);
                x[k + Q] = csub(e[k], t);
            } else {
                bfly_w32(e[k], o[k], k32, x[k], x[k + Q]);
            }
        }
    }
}

// LDS index with one padding slot per 32 complex values.  With 8-byte accesses over 64 banks this
// keeps the Stockham patterns conflict-free: contiguous runs, first-stage scatters j*R + r for
// R = 16 and 32, and runs of 16 lanes 256 apart all spread over both halves of the banks.
constexpr int PAD_SHIFT = 5;
__host__ __device__ constexpr int padi(int e) { return e + (e >> PAD_SHIFT); }
// Stride between sequences stored back to back: padded length plus one odd slot, so that lanes
// walking across sequences at a fixed position hit distinct banks.
__host__ __device__ constexpr int seq_stride(int len) { return padi(len) + 1; }

// Table of exp(-2 pi i x / 16384), x in [0, 16384), FP64-generated, in global memory.
constexpr int TW14_LOG = 14;
constexpr int TW14 = 1 << TW14_LOG;

// exp(-2 pi i y / 2^14), y in [0, 2^14), from a two-level table kept in LDS:
// twl[0..128) = w^i, twl[128..256) = w^(128 i).  Two LDS reads and one product (about 1.5 ulp).
constexpr int TWL_SIZE = 256;
__device__ __forceinline__ float2 tw_lds(const float2* twl, int y) {
    return cmul(twl[y & 127], twl[128 + ((y >> 7) & 127)]);
}
// Fills the LDS table from the global tw14 table (all threads of the workgroup take part; the
// caller synchronises before first use).
__device__ __forceinline__ void load_twl(float2* twl, const float2* __restrict__ tw14, int nthreads) {
    for (int i = threadIdx.x; i < TWL_SIZE; i += nthreads)
        twl[i] = i < 128 ? tw14[i] : tw14[((i - 128) * 128) & (TW14 - 1)];
}

// tw[r] = exp(-2 pi i r x / 2^14) for r < R from log2(R) table reads (r = 2^b) and at most
// log2(R)-1 products for the other r.
template <int R>
__device__ __forceinline__ void twiddle_powers(float2 (&tw)[R], const float2* twl, int x) {
    tw[0] = make_float2(1.f, 0.f);
#pragma unroll
    for (int b = 1; b < R; b <<= 1) tw[b] = tw_lds(twl, (b * x) & (TW14 - 1));
#pragma unroll
    for (int r = 3; r < R; ++r) {
        if ((r & (r - 1)) == 0) continue;
        const int hb = 1 << (31 - __builtin_clz(r));
        tw[r] = cmul(tw[hb], tw[r - hb]);
    }
}

// Hides an integer's value from the optimiser: per-lane address arithmetic repeated in two phases
// of a stage is otherwise computed once and kept live across them.
__device__ __forceinline__ int opaque_int(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

// Twiddle table of one Stockham stage in LDS: row q < NS holds w^(q r / (NS R)) for r < R, rows
// padded to R + 1 entries (bank spread).  Filled by all threads; the caller synchronises.
template <int R, int NS>
__device__ __forceinline__ void fill_stage_table(float2* tab, const float2* __restrict__ tw14, int nthreads) {
    for (int e = threadIdx.x; e < NS * (R + 1); e += nthreads) {
        const int q = e / (R + 1), r = e % (R + 1);
        tab[e] = r < R ? tw14[(q * (TW14 / (NS * R)) * r) & (TW14 - 1)] : make_float2(0.f, 0.f);
    }
}

// Hides a complex value from the optimiser so that quantities derived from it are computed where
// they are used instead of being hoisted and kept live.
__device__ __forceinline__ float2 opaque_f2(float2 v) {
    asm volatile("" : "+v"(v.x), "+v"(v.y));
    return v;
}

// One in-place Stockham radix-R stage over NSEQ sequences of length 2^LG in LDS.
// Reads all butterflies' inputs, barrier, writes, barrier.  TAB: twl is this stage's table
// (fill_stage_table) instead of the generic two-level table.
template <int R, int NS, int LG, int NSEQ, int SEQS, int NT, bool TAB = false>
__device__ __forceinline__ void stockham_stage(float2* buf, const float2* twl) {
    constexpr int L = 1 << LG;
    constexpr int BPS = L / R;
    constexpr int NB = NSEQ * BPS;
    constexpr int PER = (NB + NT - 1) / NT;
    float2 v[PER][R];
    const int tid = opaque_int(threadIdx.x);  // per-stage address arithmetic, never kept live
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int b = tid + i * NT;
        if ((NB % NT) == 0 || b < NB) {
            const int seq = b / BPS, j = b % BPS;
            const float2* base = buf + seq * SEQS;
            if constexpr (BPS % (1 << PAD_SHIFT) == 0) {
                // padi(j + r*BPS) = padi(j) + r*padi(BPS): one address, immediate offsets
                const float2* rb = base + padi(j);
#pragma unroll
                for (int r = 0; r < R; ++r) v[i][r] = rb[r * padi(BPS)];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) v[i][r] = base[padi(j + r * BPS)];
            }
            if constexpr (NS > 1 && TAB) {
                const float2* trow = twl + (j % NS) * (R + 1);
#pragma unroll
                for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], trow[r]);
            } else if constexpr (NS > 1) {
                float2 tw[R];
                twiddle_powers<R>(tw, twl, (j % NS) * (TW14 / (NS * R)));
#pragma unroll
                for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], tw[r]);
            }
            dft<R>(v[i]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int b = tid + i * NT;
        if ((NB % NT) == 0 || b < NB) {
            const int seq = b / BPS, j = b % BPS;
            float2* base = buf + seq * SEQS;
            const int d = (j / NS) * NS * R + (j % NS);
            if constexpr (NS % (1 << PAD_SHIFT) == 0) {
                float2* wb = base + padi(d);
#pragma unroll
                for (int r = 0; r < R; ++r) wb[r * padi(NS)] = v[i][r];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) base[padi(d + r * NS)] = v[i][r];
            }
        }
    }
    __syncthreads();
}

// First Stockham stage (span 1) of a length-L transform for one butterfly j whose inputs
// v[r] = x[j + r L/R] are already in registers: DFT, then the outputs go to LDS at j R + r.
template <int R>
__device__ __forceinline__ void stage_first_to_lds(float2 (&v)[R], float2* seqbuf, int j) {
    dft<R>(v);
#pragma unroll
    for (int r = 0; r < R; ++r) seqbuf[padi(j * R + r)] = v[r];
}

// Last Stockham stage (span L/R) for one butterfly j, read from LDS; the outputs stay in
// registers: v[r] = X[j + r L/R].
template <int R, int L>
__device__ __forceinline__ void stage_last_from_lds(float2 (&v)[R], const float2* seqbuf, int j,
                                                    const float2* twl) {
    constexpr int NS = L / R;
    if constexpr (NS % (1 << PAD_SHIFT) == 0) {
        const float2* rb = seqbuf + padi(j);
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = rb[r * padi(NS)];
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = seqbuf[padi(j + r * NS)];
    }
    float2 tw[R];
    twiddle_powers<R>(tw, twl, j * (TW14 / L));
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[r]);
    dft<R>(v);
}

// stage_last_from_lds with the twiddles of butterfly j read from a stage table
// (fill_stage_table<R, L / R>): trow = table + (j % (L / R)) * (R + 1).
template <int R, int L>
__device__ __forceinline__ void stage_last_from_lds_tab(float2 (&v)[R], const float2* seqbuf, int j,
                                                        const float2* trow) {
    constexpr int NS = L / R;
    if constexpr (NS % (1 << PAD_SHIFT) == 0) {
        const float2* rb = seqbuf + padi(j);
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = rb[r * padi(NS)];
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = seqbuf[padi(j + r * NS)];
    }
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], trow[r]);
    dft<R>(v);
}

// Radix-16 stages from span NS up to 2^LG; a final radix-32 stage absorbs a leftover factor 2.
template <int NS, int LG, int NSEQ, int SEQS, int NT>
__device__ __forceinline__ void stockham_rest(float2* buf, const float2* twl) {
    constexpr int L = 1 << LG;
    if constexpr (NS * 32 == L) {
        stockham_stage<32, NS, LG, NSEQ, SEQS, NT>(buf, twl);
    } else if constexpr (NS < L) {
        stockham_stage<16, NS, LG, NSEQ, SEQS, NT>(buf, twl);
        stockham_rest<NS * 16, LG, NSEQ, SEQS, NT>(buf, twl);
    }
}

// Forward FFT of NSEQ sequences of length 2^LG (LG <= 14) held in LDS at buf[s*SEQS + padi(e)].
// Natural order in and out.  The caller must have synchronised after filling buf; the routine
// ends with a barrier.  twl: the LDS twiddle table (tw_lds).  Plan: radix-16 stages, a leading radix-2/4/8 for LG = 2,3 (mod 4), a
// closing radix-32 for LG = 1 (mod 4), LG >= 5.
template <int LG, int NSEQ, int SEQS, int NT, bool RADIX32 = true>
__device__ __forceinline__ void lds_fft(float2* buf, const float2* twl) {
    static_assert(LG >= 1 && LG <= TW14_LOG, "FFT length out of range");
    constexpr int REM = LG % 4;
    if constexpr (REM == 1 && LG >= 5 && RADIX32) {
        stockham_rest<1, LG, NSEQ, SEQS, NT>(buf, twl);
    } else if constexpr (REM != 0) {
        stockham_stage<(1 << REM), 1, LG, NSEQ, SEQS, NT>(buf, twl);
        stockham_rest<(1 << REM), LG, NSEQ, SEQS, NT>(buf, twl);
    } else {
        stockham_rest<1, LG, NSEQ, SEQS, NT>(buf, twl);
    }
}

// exp(-2 pi i x / L) from a two-level table: lo[x mod 2^14] * hi[x >> 14].
__device__ __forceinline__ float2 tw_two_level(const float2* __restrict__ lo,
                                               const float2* __restrict__ hi, uint32_t x) {
    return cmul(lo[x & (TW14 - 1)], hi[x >> TW14_LOG]);
}

// Real-FFT split: given Z = FFT_N(x_even + i x_odd) at k and Z at (N-k) mod N, and w = exp(-2 pi i k/(2N)),
// returns 2 X[k] of the length-2N real transform (the factor 1/2 of the textbook split is left out
// here and in c2r_split; the epilogue's normalisation 1/(8N) absorbs the exact factor 8 that the
// two forward splits and the inverse split leave on the result).
__device__ __forceinline__ float2 r2c_split(float2 zk, float2 zp, float2 w) {
    const float2 zc = conjf2(zp);
    const float2 e = cadd(zk, zc);
    const float2 d = csub(zk, zc);
    const float2 t = cmul(w, d);
    return make_float2(e.x + t.y, e.y - t.x);  // e - i*t
}
// Inverse split: from P[k], P[N-k] of a Hermitian length-2N spectrum, returns
// 2 Q[k], Q[k] = FFT_N(c_even + i c_odd)[k].
__device__ __forceinline__ float2 c2r_split(float2 pk, float2 pp, float2 w) {
    const float2 pc = conjf2(pp);
    const float2 e = cadd(pk, pc);
    const float2 d = csub(pk, pc);
    const float2 t = cmul(conjf2(w), d);
    return make_float2(e.x - t.y, e.y + t.x);  // e + i*t
}
// Both splits of a partner pair k, N-k at once.  The partner's twiddle is
// exp(-2 pi i (N-k)/(2N)) = -conj(w), so its sum, difference and twiddled difference are the
// conjugate, the negated conjugate and the conjugate of the first one's: one complex product
// serves both.  r2c_split_pair returns 2 X[k], 2 X[N-k]; c2r_split_pair returns 2 Q[k], 2 Q[N-k].
__device__ __forceinline__ void r2c_split_pair(float2 zk, float2 zp, float2 w, float2& xk, float2& xp) {
    const float2 zc = conjf2(zp);
    const float2 e = cadd(zk, zc);
    const float2 t = cmul(w, csub(zk, zc));
    xk = make_float2(e.x + t.y, e.y - t.x);
    xp = make_float2(e.x - t.y, -(e.y + t.x));
}
__device__ __forceinline__ void c2r_split_pair(float2 pk, float2 pp, float2 w, float2& qk, float2& qp) {
    const float2 pc = conjf2(pp);
    const float2 e = cadd(pk, pc);
    const float2 t = cmul(conjf2(w), csub(pk, pc));
    qk = make_float2(e.x - t.y, e.y + t.x);
    qp = make_float2(e.x + t.y, t.x - e.y);
}
// 2 X[N] (Nyquist) of the length-2N real transform from Z[0], matching r2c_split's scale.
__device__ __forceinline__ float2 r2c_nyquist(float2 z0) { return make_float2(2.f * (z0.x - z0.y), 0.f); }


// Streaming stores (global_store ... nt) for data that the next pass reads only after the whole
// pass is done: spectra and output nibbles far larger than L2.  Marking them non-temporal keeps
// them from evicting the lines that concurrent workgroups still share (bit words, twiddles).
__device__ __forceinline__ void st_nt(float2* p, float2 v) {
    typedef float f2v __attribute__((ext_vector_type(2)));
    f2v x = {v.x, v.y};
    __builtin_nontemporal_store(x, reinterpret_cast<f2v*>(p));
}
__device__ __forceinline__ void st_nt_u8(uint8_t* p, uint8_t v) { __builtin_nontemporal_store(v, p); }

// corr = (wk * ws / L) mod 2, exactly: the mean term that zeroing both DC bins removes
// (calculateCorrectionFloat, PrivacyAmplification.cu:363-373, evaluated in 64-bit integers).
__device__ __forceinline__ float correction(uint32_t wk, uint32_t ws, uint32_t L) {
    const uint64_t prod = (uint64_t)wk * (uint64_t)ws;
    const uint64_t r = prod & (2ull * L - 1ull);  // L is a power of two
    return (float)((double)r / (double)L);
}

// The output bit of a pre-rounding value: round half to even, then the parity (ToBinaryArray,
// PrivacyAmplification.cu:739-783; glsl/toBinaryArray.comp:48); dist tracks max |v - rint v|.
__device__ __forceinline__ uint32_t round_bit(float v, float& dist) {
    const float r = rintf(v);
    dist = fmaxf(dist, fabsf(v - r));
    return (uint32_t)__float2int_rn(v) & 1u;
}

// Wave-level max of a non-negative float published to a device float with an atomic max
// on its bit pattern (non-negative IEEE floats order like their unsigned bit patterns).
__device__ __forceinline__ void margin_publish(float v, float* margin) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) atomicMax(reinterpret_cast<unsigned int*>(margin), __float_as_uint(v));
}

// Blockid remap so that consecutive logical ids land on the same XCD (workgroups are dealt to
// the 8 XCDs round robin): logical = (b % 8) * (G / 8) + b / 8 when G is a multiple of 8.
__device__ __forceinline__ uint32_t xcd_logical(uint32_t b, uint32_t G) {
    if ((G & 7u) != 0u) return b;
    return (b & 7u) * (G >> 3) + (b >> 3);
}

}  // namespace pa

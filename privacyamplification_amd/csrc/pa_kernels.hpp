// pa_kernels.hpp -- host-side launchers of the gfx950 kernels (internal to libpa_hip.so).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pa_geom.hpp"

namespace pa {

enum { ROWS_DYN = 0, ROWS_STATIC = 1, ROWS_SEEDPREP = 2 };
constexpr int ROWS_NT = 512;  // threads of the row pass (256-VGPR budget per lane)
constexpr int FWD_NT = 256;   // threads of the forward column pass
constexpr int TB_NT = 128;        // bit-transpose threads: 32 outputs each
constexpr int TB_SPAN_LOG = 12;   // bit-transpose outputs per workgroup (2^12 of a row of N2 = 2^13)
constexpr int POP_SLOTS = 64; // partial popcount counters per operand
constexpr int PACK_T1 = 32;   // output packer tile: rows x column pairs
constexpr int PACK_G = 256;

// Device twiddle tables: tw14[x] = exp(-2 pi i x / 2^14); lo/hi: exp(-2 pi i x / L) split as
// lo[x mod 2^14] * hi[x >> 14].
// cls: pass A's class tables, cls_table_offset(D) onwards: entry [(c NB + q) 2^BW + byte] =
// sum over the set bits b of byte of exp(-2 pi i ((BW q + b) c mod D) / D), BW = min(D, 8),
// NB = D / BW; computed in double precision and rounded once.
struct Tables {
    const float2* tw14;
    const float2* lo;
    const float2* hi;
    const float2* cls;
};
__host__ __device__ constexpr int cls_bits(int D) { return D < 8 ? D : 8; }
__host__ __device__ constexpr int cls_tables(int D) { return D / cls_bits(D); }
__host__ __device__ constexpr int cls_table_size(int D) { return D * cls_tables(D) * (1 << cls_bits(D)); }
__host__ __device__ constexpr int cls_table_offset(int D) {
    return D <= 1 ? 0 : cls_table_offset(D / 2) + cls_table_size(D / 2);
}
constexpr int CLS_TABLE_TOTAL = cls_table_offset(32);  // D = 1 .. 16

// Columns per workgroup in pass A' (tile of W x N1 complex <= 16384 values in LDS) and its threads.
// N1 = 2^10 .. 2^12 take 8192 / N1 columns (about 68 KB of LDS) and N1 = 8192 takes 2 columns one at
// a time, so that two workgroups share a compute unit.  N1 = 2^14 takes 2 columns one at a time
// (one workgroup per CU), N1 = 2^15 one column in two LDS halves.
__host__ __device__ constexpr int inv_cols_width(int lg1) {
    return lg1 >= 15 ? 1 : (lg1 >= 13 ? 2 : (lg1 >= 10 ? (1 << (13 - lg1)) : ((14 - lg1) >= 4 ? 16 : (1 << (14 - lg1)))));
}
__host__ __device__ constexpr int clamp_threads(int t) { return t < 64 ? 64 : (t > 1024 ? 1024 : t); }
__host__ __device__ constexpr int inv_cols_threads(int lg1) {
    return lg1 == 13 ? 256 : (lg1 >= 14 ? 1024 : clamp_threads(inv_cols_width(lg1) * (1 << lg1) / 16));
}
// Columns per workgroup in pass A (16; 8 when the class length M = 2048 would overflow LDS).
__host__ __device__ constexpr int fwd_cols_width(int lgm) { return lgm >= 11 ? 8 : 16; }
// Threads of the generic pass A: 1024 for classes of 1024 and 2048 points, whose 128 KB of columns
// leave room for one workgroup per CU, so that it brings 16 waves instead of 4.
__host__ __device__ constexpr int fwd_cols_threads(int lgm) { return lgm >= 10 ? 1024 : FWD_NT; }
// Bytes of the pass-A' output: nibbles per column pair (N/2), or 2-bit crumbs per column (N) when
// a workgroup holds a single column.
__host__ __device__ constexpr uint64_t nib_bytes(uint32_t N, uint32_t W) { return (W == 1 ? (uint64_t)N : N / 2) + 64; }
__host__ __device__ constexpr int small_threads(int lgn) { return clamp_threads(2 * (1 << lgn) / 16); }

// A staging range: `bytes` from src to dst, either side device memory or pinned, device-mapped
// host memory.
struct CopyDesc {
    const uint8_t* src;
    uint8_t* dst;
    uint64_t bytes;
};
// Copies every range of d[0 .. count) in one launch; max_bytes = the longest range.
hipError_t launch_copies(const CopyDesc* d, uint32_t count, uint64_t max_bytes, hipStream_t s);
hipError_t launch_zero_u32(uint32_t* p, uint32_t count, hipStream_t s);
// bits2/tb2 (and tb2/Y2) non-null: the seed is transformed in the same launch (grid.y = 2),
// its popcount into count + POP_SLOTS.
hipError_t launch_transpose(const uint32_t* bits, uint32_t* tb, uint32_t* count, const Geom& g,
                            int is_seed, hipStream_t s, const uint32_t* bits2 = nullptr, uint32_t* tb2 = nullptr);
hipError_t launch_fwd_cols(const uint32_t* tb, float2* Y, const Geom& g, const Tables& t, hipStream_t s,
                           const uint32_t* tb2 = nullptr, float2* Y2 = nullptr);
// ROWS_SEEDPREP: S <- split spectrum of the seed rows held in S (in place, N+1 values).
// ROWS_STATIC:   R' from the key rows in Yk and the split seed spectrum S.
// ROWS_DYN:      R' from the key rows in Yk and the seed rows (pass-A output) in S.
// R' goes to Rt in tiles when rt_tiled(g.lg1) (N complex values), else to Yk in place.
// Persistent: at most `cus` workgroups (one per compute unit) walk the row pairs.
hipError_t launch_rows(int mode, float2* Yk, float2* S, float2* Rt, const Geom& g, const Tables& t, int cus,
                       hipStream_t s);
// R: pass B's output (the tile buffer when rt_tiled(g.lg1)).
hipError_t launch_inv_cols(const float2* R, uint8_t* nib, const Geom& g, const uint32_t* wk,
                           const uint32_t* ws, const Tables& t, float* margin, hipStream_t s);
// zero_counts != null: the launch also clears those 2 * POP_SLOTS popcount partials.
hipError_t launch_pack(const uint8_t* nib, const uint32_t* key, uint32_t* out, const Geom& g, hipStream_t s,
                       uint32_t* zero_counts = nullptr);
hipError_t launch_small(const uint32_t* key, const uint32_t* seed, const float2* sspec,
                        const uint32_t* ws_ptr, uint32_t* out, const Geom& g, const Tables& t,
                        float* margin, float2* sspec_out, uint32_t* ws_out, hipStream_t s);

// Batched forms: `count` blocks of one transform length in one launch each (grid z = block),
// block b's geometry and buffers in bt[b] (device memory); g0 is any block's geometry.
uint32_t pack_tiles(const Geom& g);  // grid.y the pack stage needs for this block's window
hipError_t launch_batch_zero(const BlockArgs* bt, uint32_t count, hipStream_t s);
hipError_t launch_batch_transpose(const BlockArgs* bt, uint32_t count, const Geom& g0, int is_seed, hipStream_t s);
// per: column groups each workgroup of the persistent pass A (N1 = 2^9 .. 2^13) walks, a power of
// two up to 32; 0 picks 16, halved while the launch would have under 512 workgroups.
hipError_t launch_batch_fwd_cols(const BlockArgs* bt, uint32_t count, const Geom& g0, const Tables& t, int is_seed,
                                 uint32_t per, hipStream_t s);
constexpr uint32_t FWD_PER_MAX = 32;
// grid_x workgroups per block walk that block's row pairs (0 = one per row pair).
hipError_t launch_batch_rows(int mode, const BlockArgs* bt, uint32_t count, const Geom& g0, const Tables& t,
                             uint32_t grid_x, hipStream_t s);
hipError_t launch_batch_inv_cols(const BlockArgs* bt, uint32_t count, const Geom& g0, const Tables& t, hipStream_t s);
hipError_t launch_batch_pack(const BlockArgs* bt, uint32_t count, const Geom& g0, uint32_t tiles_y, hipStream_t s);
hipError_t launch_batch_small(const BlockArgs* bt, uint32_t count, const Geom& g0, const Tables& t, hipStream_t s);

// Sub-block split (tiles of b x b through the batched engine, XOR-accumulated per band).
hipError_t launch_tile_inputs(const uint32_t* key, const uint32_t* seed, uint32_t n, uint32_t h, uint32_t b,
                              uint32_t nbands, uint32_t j0, uint32_t tiles, uint32_t* tkeys, uint32_t* tseeds,
                              hipStream_t s);
hipError_t launch_tile_xor(uint32_t* acc, const uint32_t* tout, uint32_t nbands, uint32_t wb, uint32_t cols,
                           int first, hipStream_t s);
hipError_t launch_split_finish(const uint32_t* acc, const uint32_t* key, uint32_t* out, uint32_t m, uint32_t h,
                               uint32_t do_xor, uint32_t reverse_endian, hipStream_t s);

// Self-test kernels (pa_selftest.hip): production device functions on closed-form inputs.
hipError_t launch_st_correction(const uint32_t* wk, const uint32_t* ws, uint32_t L, float* out, uint32_t count,
                                hipStream_t s);
hipError_t launch_st_product(const float2* a, const float2* b, float r, float2* prod, const float2* z, float2* zz,
                             uint32_t N, hipStream_t s);
hipError_t launch_st_nibbles(uint8_t* nib, const Geom& g, float corr, float invN, hipStream_t s);

}  // namespace pa

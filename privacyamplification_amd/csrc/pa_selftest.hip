// pa_selftest.hip -- device halves of the engine's self-tests (pa_selftest): small kernels that run
// the production device functions (pa_device.hpp) on the closed-form inputs of the reference's
// unit tests (PrivacyAmplification.cu:299-737).  The rest of each test drives the production
// kernels themselves through their launchers (pa_engine.cpp).
#include <hip/hip_runtime.h>

#include "pa_device.hpp"
#include "pa_geom.hpp"
#include "pa_kernels.hpp"

namespace pa {

// out[i] = correction(wk[i], ws[i], L): the correction term exactly as pass A' computes it
// (unitTestCalculateCorrectionFloat's inputs, PrivacyAmplification.cu:299-361).
__global__ __launch_bounds__(256) void k_st_correction(const uint32_t* __restrict__ wk, const uint32_t* __restrict__ ws,
                                                       uint32_t L, float* __restrict__ out, uint32_t count) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < count) out[i] = correction(wk[i], ws[i], L);
}

// The spectral product as the row pass forms it: the complex product of the key's and the seed's
// values (ElementWiseProduct, PrivacyAmplification.cu:465-481, with its 1/r pre-scaling applied to
// both operands as the reference does), and the round trip of the real-FFT splits that surround it:
// c2r_split_pair(r2c_split_pair(z)) = 4 z for every partner pair (k, N-k).
__global__ __launch_bounds__(256) void k_st_product(const float2* __restrict__ a, const float2* __restrict__ b,
                                                    float r, float2* __restrict__ prod, const float2* __restrict__ z,
                                                    float2* __restrict__ zz, uint32_t N) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < N) {
        const float inv = 1.0f / r;
        prod[i] = cmul(cscale(a[i], inv), cscale(b[i], inv));
    }
    if (i >= 1 && i < N / 2) {  // pair (i, N - i), split twiddle w = exp(-2 pi i i / (2N))
        const double ang = -3.14159265358979323846 * (double)i / (double)N;
        const float2 w = make_float2((float)cos(ang), (float)sin(ang));
        float2 xk, xp, qk, qp;
        r2c_split_pair(z[i], z[N - i], w, xk, xp);
        c2r_split_pair(xk, xp, w, qk, qp);
        zz[i] = qk;
        zz[N - i] = qp;
    }
}

// Pass A' output for the ToBinaryArray closed form (PrivacyAmplification.cu:669-737): the
// pre-rounding value of conv index p is v[p] = bit 31 - (p mod 32) of the integer p / 32 (0 or 1);
// each nibble (or crumb) holds round_bit(v * invN + corr) of its positions, in pass A''s layout:
// nibble nibT[(t2 / 2) N1 + t1] = bits of 2(t2 + N2 t1), +1 for columns t2, t2 + 1 (t2 even), MSB
// first; crumb nibT[t2 N1 + t1] = bits of 2(t2 + N2 t1), +1.
__device__ __forceinline__ float st_pattern(uint32_t p) { return (float)(((p >> 5) >> (31u - (p & 31u))) & 1u); }
__global__ __launch_bounds__(256) void k_st_nibbles(uint8_t* __restrict__ nibT, Geom g, float corr, float invN) {
    const uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    float dist = 0.f;
    if (g.W == 1) {
        if (e >= (uint64_t)g.N) return;
        const uint32_t t2 = (uint32_t)(e / g.N1), t1 = (uint32_t)(e % g.N1);
        const uint32_t i0 = 2u * (t2 + g.N2 * t1);
        const uint32_t b0 = round_bit(st_pattern(i0) * invN + corr, dist);
        const uint32_t b1 = round_bit(st_pattern(i0 + 1u) * invN + corr, dist);
        nibT[e] = (uint8_t)((b0 << 1) | b1);
    } else {
        if (e >= (uint64_t)g.N / 2) return;
        const uint32_t pair = (uint32_t)(e / g.N1), t1 = (uint32_t)(e % g.N1);
        const uint32_t i0 = 2u * (2u * pair + g.N2 * t1), i1 = i0 + 2u;
        const uint32_t b0 = round_bit(st_pattern(i0) * invN + corr, dist);
        const uint32_t b1 = round_bit(st_pattern(i0 + 1u) * invN + corr, dist);
        const uint32_t b2 = round_bit(st_pattern(i1) * invN + corr, dist);
        const uint32_t b3 = round_bit(st_pattern(i1 + 1u) * invN + corr, dist);
        nibT[e] = (uint8_t)((b0 << 3) | (b1 << 2) | (b2 << 1) | b3);
    }
}

hipError_t launch_st_correction(const uint32_t* wk, const uint32_t* ws, uint32_t L, float* out, uint32_t count,
                                hipStream_t s) {
    k_st_correction<<<(count + 255u) / 256u, 256, 0, s>>>(wk, ws, L, out, count);
    return hipGetLastError();
}
hipError_t launch_st_product(const float2* a, const float2* b, float r, float2* prod, const float2* z, float2* zz,
                             uint32_t N, hipStream_t s) {
    k_st_product<<<(N + 255u) / 256u, 256, 0, s>>>(a, b, r, prod, z, zz, N);
    return hipGetLastError();
}
hipError_t launch_st_nibbles(uint8_t* nib, const Geom& g, float corr, float invN, hipStream_t s) {
    const uint64_t cnt = g.W == 1 ? (uint64_t)g.N : (uint64_t)g.N / 2;
    k_st_nibbles<<<(uint32_t)((cnt + 255u) / 256u), 256, 0, s>>>(nib, g, corr, invN);
    return hipGetLastError();
}

}  // namespace pa

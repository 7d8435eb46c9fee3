// Tip5 permutation for one sponge per lane on gfx950.
//
// Follows the reference permutation (src/hash/tip5.cpp:92-203): 5 rounds of
// S-box -> circulant MDS -> round constants; S-box = split-and-lookup on lanes
// 0..3 and x^7 on lanes 4..15.  The arithmetic is re-derived for VALU:
//
//  * the state is kept in Montgomery form x_M = x * 2^64 mod p (the form the
//    reference's split-and-lookup is defined on): the reference forms the
//    canonical Montgomery value montyred(v * R^2) = v * 2^64 mod p, looks up each
//    byte and maps back with montyred(0, o) = o * 2^-64, so in Montgomery form
//    the S-box is the byte lookup of the canonical value alone, and x^7 uses
//    Montgomery products (montyred is cheaper than the 2^64 = 2^32 - 1 fold).
//    Callers convert absorbed elements in (to_mont) and digests out
//    (montyred(y, 0), canonical).  The 256-byte table sits in LDS.
//  * MDS: the 16x16 circulant has 16-bit entries, so each state word is split
//    into 32-bit halves and both halves are multiplied in f64 (exact: every
//    partial sum is below 2^52), using the x^8 - 1 / x^8 + 1 split of the
//    cyclic convolution.  The two 52-bit results are recombined modulo p with
//    integer ops.
//  * state words stay "lazy" (any u64 congruent mod p) inside the permutation;
//    callers canonicalise what they store.
#pragma once

#include "gl64.hpp"

namespace tip5 {

using gl::u32;
using gl::u64;

constexpr int STATE = 16;
constexpr int RATE = 10;
constexpr int ROUNDS = 5;
constexpr int DIGEST = 5;

// Round constants of the Tip5 specification (src/hash/tip5.cpp:48-79).
constexpr u64 kRoundConstants[ROUNDS * STATE] = {
    0xbd2a3deb61ab60deull, 0xea7df21ad9547ed2ull, 0x900b3677a1de063full, 0x1b46887e876c8677ull,
    0xd364d977889cfb97ull, 0xdc8dfac843699f02ull, 0x375c405d7190db58ull, 0x27924006d2b0d4b1ull,
    0x78dd1172d483cd38ull, 0x3346c66244882a56ull, 0xb0249b279f498aa5ull, 0x94cd51be79338d4dull,
    0xb0e0dc7052c5b218ull, 0xf8dcc4d248adad95ull, 0x68e3c635fec868b7ull, 0xd7d06b3ffb6b0d8cull,
    0xf3500dea20ef032aull, 0x4865bf175bba5803ull, 0xd5f7fe3027287a27ull, 0xa57333f44e193412ull,
    0x8726e153a977eae2ull, 0x3014a98463fc191bull, 0xba145461af39b212ull, 0x03ab70105933202full,
    0x3d90b7eebfcf71e5ull, 0x386322b1cc520bfdull, 0x27c2c8daf774f675ull, 0x4fcb83f50309bc6aull,
    0x5e6d5ce8275f3cb3ull, 0xecc2f6592c8f905cull, 0x837f532461e609b4ull, 0xb2b1f6b95c92c93cull,
    0xc0027af556411dc1ull, 0x16e18c885fc2a26cull, 0x8880ef183d9f2bf3ull, 0xb2930bdb5ca88c45ull,
    0x9c2ec8322e1c1553ull, 0xe5b05eaf3220a674ull, 0xa49cc6ae4b861c4eull, 0x11708e0aeb86ebd7ull,
    0xc09de92bbc3902e0ull, 0x929b3c79516bcbc1ull, 0xe006e5bf738f27d1ull, 0x2d9e1ec0eac8ea38ull,
    0x0984d8d94bf937c5ull, 0x4959273c220e6747ull, 0xfe1d934207e796faull, 0x2b9b9298f2f6dd73ull,
    0x07a1f5a67d6e3a41ull, 0x4407593ee73743d9ull, 0x9f054720ef802e59ull, 0x78d4b711336e6aa6ull,
    0xadc638aef3c8b228ull, 0xa4d6d3e86afb2114ull, 0x9d4808e725531968ull, 0x369804df3866d0efull,
    0xe6dbd9a9d2215024ull, 0x8ed22ca212ee85b2ull, 0x397bb882fcd23eb6ull, 0xeb8f8786d7277531ull,
    0x9999d4cdaff543b5ull, 0xf382a61217f192d6ull, 0x49c37260b026adc1ull, 0x3ff8918ce35c1019ull,
    0x2e7df8b76080bd07ull, 0xf5dbac250b8a28b9ull, 0x853c3727ae9da4ccull, 0xb2f1f5f3d9e5a26dull,
    0x3fce22012d337847ull, 0x6b5a3e6db7eee347ull, 0x171582cd59dde50dull, 0xc0c0b3095ee62a8aull,
    0x665b25c6f6a203d2ull, 0x3099aed93b6ae69full, 0x801df6092be69c38ull, 0x8066ad0cdfff43cdull,
    0x8af9d44a5f4fdc6bull, 0xd80219cd97c0d762ull, 0x10c9ceba14148ebbull, 0x539bd4c3f2f24474ull,
};

struct RoundConstantsM {
    u64 v[ROUNDS * STATE];
};
constexpr RoundConstantsM mont_round_constants() {
    RoundConstantsM t{};
    for (int i = 0; i < ROUNDS * STATE; i++)
        t.v[i] = (u64)(((unsigned __int128)kRoundConstants[i] * gl::EPS) % gl::P);
    return t;
}
// (the round constants in Montgomery form: c * 2^64 mod p, 2^64 = 2^32 - 1 mod p)

// 1 in Montgomery form (2^64 mod p)
constexpr u64 kOneM = gl::EPS;

// First column of the circulant MDS matrix: SHA-256("Tip5") as 16 LE u16:
// 61402 1108 28750 33823 7454 43244 53865 12034 56951 27521 41351 40901 12021
// 59689 26798 17845 (folded into kMdsP / kMdsQ below).

// Fill a 256-byte LDS table with the S-box ((x+1)^3 mod 257) - 1.
__device__ __forceinline__ void load_lut(uint8_t* lut) {
    for (u32 x = threadIdx.x; x < 256; x += blockDim.x) {
        u32 t = x + 1;
        lut[x] = (uint8_t)(((t * t % 257) * t % 257) - 1);
    }
}

// S-box of lanes 0..3 on a Montgomery-form value: bytes of the canonical
// Montgomery value through the table; the result is the Montgomery form of the
// reference's output (tip5.cpp split_and_lookup).
__device__ __forceinline__ u64 split_and_lookup_m(u64 v, const uint8_t* lut) {
    u64 m = gl::canon(v);
    u32 lo = (u32)m, hi = (u32)(m >> 32);
    u32 olo = (u32)lut[lo & 0xff] | ((u32)lut[(lo >> 8) & 0xff] << 8) |
              ((u32)lut[(lo >> 16) & 0xff] << 16) | ((u32)lut[lo >> 24] << 24);
    u32 ohi = (u32)lut[hi & 0xff] | ((u32)lut[(hi >> 8) & 0xff] << 8) |
              ((u32)lut[(hi >> 16) & 0xff] << 16) | ((u32)lut[hi >> 24] << 24);
    return ((u64)ohi << 32) | olo;
}

__device__ __forceinline__ u64 pow7_m(u64 x) {
    u64 x2 = gl::mont_sqr(x);
    u64 x4 = gl::mont_sqr(x2);
    u64 x3 = gl::mont_mul(x, x2);
    return gl::mont_mul(x3, x4);
}

// integer-valued double a >= 0 and bias = 2^52 + c with a + c < 2^52 -> a + c
__device__ __forceinline__ u64 f2u52(double a, double bias) {
    return (u64)__double_as_longlong(a + bias) & 0x000FFFFFFFFFFFFFull;
}

// Round constants as f64 biases for the MDS recombination: adding the 32-bit
// halves c0, c1 of the Montgomery-form constant to the exact MDS outputs
// (L + c0) + (H + c1) * 2^32 adds the constant itself; 2^52 + c is exact in f64
// and L + c0, H + c1 stay below 2^52.
struct RcBias {
    double v[ROUNDS * STATE * 2];
};
constexpr RcBias rc_bias_table() {
    RcBias t{};
    const RoundConstantsM m = mont_round_constants();
    for (int i = 0; i < ROUNDS * STATE; i++) {
        u64 c = m.v[i];
        t.v[2 * i] = 4503599627370496.0 + (double)(u32)c;
        t.v[2 * i + 1] = 4503599627370496.0 + (double)(u32)(c >> 32);
    }
    return t;
}
__constant__ RcBias kRcBias = rc_bias_table();

// Circulant MDS y_r = sum_c M[(r - c) & 15] s_c on the two 32-bit halves of the
// state, each via x^16 - 1 = (x^8 - 1)(x^8 + 1): a cyclic 8-term product of
// (s_lo + s_hi) with kMdsP = (m_lo + m_hi)/2 and a negacyclic one of
// (s_lo - s_hi) with kMdsQ = (m_lo - m_hi)/2 give A and B with y_lo = A + B,
// y_hi = A - B.  Every partial sum is a multiple of 1/2 below 2^52 in
// magnitude, so all f64 arithmetic is exact (128 FMAs per half instead of 256).
constexpr double kMdsP[8] = {(61402 + 56951) / 2.0, (1108 + 27521) / 2.0, (28750 + 41351) / 2.0,
                             (33823 + 40901) / 2.0, (7454 + 12021) / 2.0, (43244 + 59689) / 2.0,
                             (53865 + 26798) / 2.0, (12034 + 17845) / 2.0};
constexpr double kMdsQ[8] = {(61402 - 56951) / 2.0, (1108 - 27521) / 2.0, (28750 - 41351) / 2.0,
                             (33823 - 40901) / 2.0, (7454 - 12021) / 2.0, (43244 - 59689) / 2.0,
                             (53865 - 26798) / 2.0, (12034 - 17845) / 2.0};

__device__ __forceinline__ void mds_half(const double x[STATE], double y[STATE]) {
    double sp[8], sm[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        sp[c] = x[c] + x[c + 8];
        sm[c] = x[c] - x[c + 8];
    }
#pragma unroll
    for (int r = 0; r < 8; r++) {
        double A = 0.0, B = 0.0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            A = __fma_rn(kMdsP[(r - c) & 7], sp[c], A);
            B = __fma_rn(r >= c ? kMdsQ[r - c] : -kMdsQ[r - c + 8], sm[c], B);
        }
        y[r] = A + B;
        y[r + 8] = A - B;
    }
}

// MDS followed by the round constants of round `round` (any u64 in, < 2^64 out).
__device__ __forceinline__ void mds_rc(u64 s[STATE], int round) {
    double lo[STATE], hi[STATE], ylo[STATE], yhi[STATE];
#pragma unroll
    for (int c = 0; c < STATE; c++) {
        lo[c] = (double)(u32)s[c];
        hi[c] = (double)(u32)(s[c] >> 32);
    }
    mds_half(lo, ylo);
    mds_half(hi, yhi);
    const double* bias = kRcBias.v + round * STATE * 2;
#pragma unroll
    for (int r = 0; r < STATE; r++) {  // (L + c0) + (H + c1) * 2^32
        s[r] = gl::lin52(f2u52(ylo[r], bias[2 * r]), f2u52(yhi[r], bias[2 * r + 1]));
    }
}

// The permutation on a Montgomery-form state (any u64 words congruent to x_M).
__device__ __forceinline__ void permutation_m(u64 s[STATE], const uint8_t* lut) {
#pragma unroll 1
    for (int r = 0; r < ROUNDS; r++) {
#pragma unroll
        for (int i = 0; i < 4; i++) s[i] = split_and_lookup_m(s[i], lut);
#pragma unroll
        for (int i = 4; i < STATE; i++) s[i] = pow7_m(s[i]);
        mds_rc(s, r);
    }
}

// Montgomery form <-> canonical
__device__ __forceinline__ u64 to_m(u64 x) { return gl::to_mont(x); }
__device__ __forceinline__ u64 from_m(u64 y) { return gl::montyred(y, 0); }

}  // namespace tip5

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
#include "tip5_constants.hpp"

namespace tip5 {

using gl::u32;
using gl::u64;

constexpr int STATE = 16;
constexpr int RATE = 10;
constexpr int ROUNDS = 5;
constexpr int DIGEST = 5;

// Round constants of the Tip5 specification: tip5_constants.hpp.
using tip5c::kRoundConstants;

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

// First column of the circulant MDS matrix: tip5c::kMdsFirstColumn (SHA-256
// of "Tip5" as 16 LE u16), folded into kMdsP / kMdsQ below.

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
struct MdsHalves {
    double v[8];
};
constexpr MdsHalves mds_halves(int sign) {
    MdsHalves t{};
    for (int i = 0; i < 8; i++)
        t.v[i] = ((double)tip5c::kMdsFirstColumn[i] + sign * (double)tip5c::kMdsFirstColumn[i + 8]) / 2.0;
    return t;
}
constexpr MdsHalves kMdsP = mds_halves(1);
constexpr MdsHalves kMdsQ = mds_halves(-1);

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
            A = __fma_rn(kMdsP.v[(r - c) & 7], sp[c], A);
            B = __fma_rn(r >= c ? kMdsQ.v[r - c] : -kMdsQ.v[r - c + 8], sm[c], B);
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

// Round r of the permutation (tip5.cpp:115-125: S-box layer, MDS, round
// constants) on a Montgomery-form state (any u64 words congruent to x_M).
__device__ __forceinline__ void round_m(u64 s[STATE], int r, const uint8_t* lut) {
#pragma unroll
    for (int i = 0; i < 4; i++) s[i] = split_and_lookup_m(s[i], lut);
#pragma unroll
    for (int i = 4; i < STATE; i++) s[i] = pow7_m(s[i]);
    mds_rc(s, r);
}

// The permutation on a Montgomery-form state.
__device__ __forceinline__ void permutation_m(u64 s[STATE], const uint8_t* lut) {
#pragma unroll 1
    for (int r = 0; r < ROUNDS; r++) round_m(s, r, lut);
}

// The permutation spread over a group of 16 lanes (16-aligned within the wave),
// for latency-bound work such as the top levels of a Merkle tree: lane w of the
// group holds state word w (Montgomery form).  The S-box acts lane-wise; the
// circulant MDS y_w = sum_c M[(w - c) & 15] s_c gathers the group's words with
// lane shuffles and sums each 32-bit half in f64, exact because sum_c M_c <
// 2^20 keeps both sums below 2^52.  Every lane of the group must take part.
struct MdsColumnF64 {
    double v[STATE];
};
constexpr MdsColumnF64 mds_column_f64() {
    MdsColumnF64 t{};
    for (int i = 0; i < STATE; i++) t.v[i] = (double)tip5c::kMdsFirstColumn[i];
    return t;
}
__constant__ MdsColumnF64 kMdsColumnF64 = mds_column_f64();

__device__ __forceinline__ u64 permutation_group16(u64 s, u32 w, const uint8_t* lut) {
    double m[STATE];
#pragma unroll
    for (int c = 0; c < STATE; c++) m[c] = kMdsColumnF64.v[(w - (u32)c) & 15];
#pragma unroll 1
    for (int r = 0; r < ROUNDS; r++) {
        s = w < 4 ? split_and_lookup_m(s, lut) : pow7_m(s);
        const u32 lo = (u32)s, hi = (u32)(s >> 32);
        double l0 = 0.0, l1 = 0.0, h0 = 0.0, h1 = 0.0;  // two chains per half
#pragma unroll
        for (int c = 0; c < STATE; c += 2) {
            l0 = __fma_rn(m[c], (double)__shfl(lo, c, 16), l0);
            h0 = __fma_rn(m[c], (double)__shfl(hi, c, 16), h0);
            l1 = __fma_rn(m[c + 1], (double)__shfl(lo, c + 1, 16), l1);
            h1 = __fma_rn(m[c + 1], (double)__shfl(hi, c + 1, 16), h1);
        }
        const double* bias = kRcBias.v + (r * STATE + w) * 2;
        s = gl::lin52(f2u52(l0 + l1, bias[0]), f2u52(h0 + h1, bias[1]));
    }
    return s;
}

// Montgomery form <-> canonical
__device__ __forceinline__ u64 to_m(u64 x) { return gl::to_mont(x); }
__device__ __forceinline__ u64 from_m(u64 y) { return gl::montyred(y, 0); }

}  // namespace tip5

// Goldilocks field p = 2^64 - 2^32 + 1 for gfx950 VALU (and the host).
//
// Values are u64.  "canonical" means in [0, p) -- the representation of the
// reference's BFieldElement (include/types/b_field_element.hpp:31-41) and of
// every buffer crossing the C ABI.  Inside kernels some paths keep values
// "lazy": any u64 congruent to the field element, made canonical before they
// are stored.  Every function says which it takes and returns.
//
// The 64x64 product is built from four 32x32->64 multiply-adds
// (v_mad_u64_u32); the 128-bit result is folded with 2^64 = 2^32 - 1 and
// 2^96 = -1 (mod p).  Multiplication by 2^k (the radix-64 twiddles: the
// reference's w_64 = 7^((p-1)/64) equals 2^39) is a shift plus the same fold.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gl {

typedef uint64_t u64;
typedef uint32_t u32;

constexpr u64 P = 0xFFFFFFFF00000001ull;
constexpr u64 EPS = 0xFFFFFFFFull;  // 2^64 mod p

__host__ __device__ __forceinline__ u64 mk64(u32 lo, u32 hi) { return ((u64)hi << 32) | lo; }

// Starts a carry chain for the device compiler: the value is pinned in a VGPR
// as it stands, so the chain that produced it cannot be merged into the next.
// (The gfx950 compiler of ROCm 7.2 folds a final subtract-with-borrow of 0 into
// a following add-with-carry of 0, as in canon(montyred(..)) or canon(sub(..)),
// and then selects on a wrong carry; the device check
// tests/native/gl64_device_check.hip shows it.)  canon() pins its input.  No
// instruction is emitted; a no-op on the host.
__host__ __device__ __forceinline__ void opaque(u32& x) {
#ifdef __HIP_DEVICE_COMPILE__
    asm volatile("" : "+v"(x));
#else
    (void)x;
#endif
}

// Explicit 32-bit carry chains (v_add_co/v_addc, v_sub_co/v_subb on gfx950):
// full-rate VALU ops, no 64-bit compares.  The same code serves the host.

// any u64 -> canonical, for an input whose producer is known to compile
// correctly in front of this chain (a 64-bit fold, a load)
__host__ __device__ __forceinline__ u64 canon_raw(u64 x) {
    u32 k, c;
    u32 t0 = __builtin_addc((u32)x, 0xFFFFFFFFu, 0u, &k);  // x + EPS carries out iff x >= p
    u32 t1 = __builtin_addc((u32)(x >> 32), 0u, k, &c);
    return c ? mk64(t0, t1) : x;
}

// any u64 -> canonical.  The input is pinned (opaque) so that the carry chain
// that produced it is never merged into this one.
__host__ __device__ __forceinline__ u64 canon(u64 x) {
    u32 x0 = (u32)x, x1 = (u32)(x >> 32);
    opaque(x0);
    opaque(x1);
    return canon_raw(mk64(x0, x1));
}

// a, b canonical -> canonical
__host__ __device__ __forceinline__ u64 add(u64 a, u64 b) {
    u32 c0, c1, c2, c3;
    u32 s0 = __builtin_addc((u32)a, (u32)b, 0u, &c0);
    u32 s1 = __builtin_addc((u32)(a >> 32), (u32)(b >> 32), c0, &c1);
    u32 t0 = __builtin_addc(s0, 0xFFFFFFFFu, 0u, &c2);
    u32 t1 = __builtin_addc(s1, 0u, c2, &c3);
    return (c1 | c3) ? mk64(t0, t1) : mk64(s0, s1);
}

// a, b canonical -> canonical
__host__ __device__ __forceinline__ u64 sub(u64 a, u64 b) {
    u32 b0, b1, b2, b3;
    u32 d0 = __builtin_subc((u32)a, (u32)b, 0u, &b0);
    u32 d1 = __builtin_subc((u32)(a >> 32), (u32)(b >> 32), b0, &b1);
    u32 e0 = __builtin_subc(d0, 0xFFFFFFFFu, 0u, &b2);   // borrow: add p = subtract EPS
    u32 e1 = __builtin_subc(d1, 0u, b2, &b3);
    return b1 ? mk64(e0, e1) : mk64(d0, d1);
}

// a any u64, b canonical -> lazy (any u64).  Used to add round constants.
__host__ __device__ __forceinline__ u64 add_lazy(u64 a, u64 b) {
    u32 c0, c1, c2, c3;
    u32 s0 = __builtin_addc((u32)a, (u32)b, 0u, &c0);
    u32 s1 = __builtin_addc((u32)(a >> 32), (u32)(b >> 32), c0, &c1);
    // carry: subtract p = add EPS; a + b < 2^65 - 2^32 so this cannot carry again
    u32 t0 = __builtin_addc(s0, c1 ? 0xFFFFFFFFu : 0u, 0u, &c2);
    u32 t1 = __builtin_addc(s1, 0u, c2, &c3);
    return mk64(t0, t1);
}

// 128-bit value (hi:lo) -> lazy u64 congruent to it:
// lo + hl*2^64 + hh*2^96 = lo + hl*(2^32 - 1) - hh  (mod p).
// (Measured on gfx950: the compiler's compare/select form of this fold beats an
// explicit carry chain, whose SGPR carries add wait states.)
__host__ __device__ __forceinline__ u64 fold128(u64 lo, u64 hi) {
    u32 hh = (u32)(hi >> 32), hl = (u32)hi;
    u64 t = lo - hh;
    if (lo < hh) t -= EPS;                 // borrow: add p
    u64 u = ((u64)hl << 32) - hl;          // hl * (2^32 - 1) < 2^64
    u64 r = t + u;
    if (r < u) r += EPS;                   // carry: subtract p
    return r;
}

// a canonical -> canonical
__host__ __device__ __forceinline__ u64 neg(u64 a) { return a ? P - a : 0; }

// 64x64 -> 128 from four independent 32x32 products (v_mad_u64_u32 on the
// device): the cross sum S = a1*b0 + a0*b1 is kept as 64 bits plus a carry c,
// then lo = p00 + (S << 32), hi = p11 + (S >> 32) + (c << 32) + carry(lo).
// No 32-bit half is widened into a fresh register pair (those moves dominated
// the chained form p00 -> a0*b1 + (p00 >> 32) -> ...).  Host code runs the
// same function in the native check against unsigned __int128.
__host__ __device__ __forceinline__ void mul_wide_4mad(u64 a, u64 b, u64& lo, u64& hi) {
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 p00 = (u64)a0 * b0;
    u64 q = (u64)a1 * b1;
    u64 x = (u64)a1 * b0, y = (u64)a0 * b1;
    u32 k0, c, k1, k2, k3, k4;
    u32 S0 = __builtin_addc((u32)x, (u32)y, 0u, &k0);  // cross sum in words: the 64-bit
    u32 S1 = __builtin_addc((u32)(x >> 32), (u32)(y >> 32), k0, &c);  // add + compare costs more
    u32 l1 = __builtin_addc((u32)(p00 >> 32), S0, 0u, &k1);
    u32 h0 = __builtin_addc((u32)q, S1, k1, &k2);
    u32 h1 = __builtin_addc((u32)(q >> 32), 0u, k2, &k3);
    h1 = __builtin_addc(h1, 0u, c, &k4);
    lo = mk64((u32)p00, l1);
    hi = mk64(h0, h1);
}

// a^2 -> 128 bits from three products: the cross term a0*a1 is doubled with
// word shifts (bit 63 becomes the carry into the high word).
__host__ __device__ __forceinline__ void sqr_wide(u64 a, u64& lo, u64& hi) {
    u32 a0 = (u32)a, a1 = (u32)(a >> 32);
    u64 p00 = (u64)a0 * a0;
    u64 q = (u64)a1 * a1;
    u64 x = (u64)a0 * a1;
    u32 x0 = (u32)x, x1 = (u32)(x >> 32);
    u32 s0 = x0 << 1, s1 = (x1 << 1) | (x0 >> 31), c = x1 >> 31;  // 2x = c*2^64 + (s1:s0)
    u32 k1, k2, k3, k4;
    u32 l1 = __builtin_addc((u32)(p00 >> 32), s0, 0u, &k1);
    u32 h0 = __builtin_addc((u32)q, s1, k1, &k2);
    u32 h1 = __builtin_addc((u32)(q >> 32), c, k2, &k3);
    (void)k3;
    (void)k4;
    lo = mk64((u32)p00, l1);
    hi = mk64(h0, h1);
}

__host__ __device__ __forceinline__ void mul_wide(u64 a, u64 b, u64& lo, u64& hi) {
#ifdef __HIP_DEVICE_COMPILE__
    mul_wide_4mad(a, b, lo, hi);
#else
    unsigned __int128 pr = (unsigned __int128)a * b;
    lo = (u64)pr;
    hi = (u64)(pr >> 64);
#endif
}

// Montgomery reduction for R = 2^64: x = hi*2^64 + lo -> x * 2^-64 mod p, for
// any 128-bit x (the reference's montyred, b_field_element.hpp:103).  With
// m = lo * (1 + 2^32) mod 2^64 (lo + (lo << 32), overflow e), x + m*p is a
// multiple of 2^64 and (x + m*p) / 2^64 = hi - b with b = m - (m >> 32) - e;
// a borrow is corrected by adding p (subtracting 2^32 - 1 mod 2^64).  The
// result is < 2^64 and canonical when hi < p.
// In 32-bit words: a = (lo0, lo1 + lo0) with e the carry of the high word,
// b = a - a1 - e, r = hi - b, then r - (2^32 - 1) on a borrow.
__host__ __device__ __forceinline__ u64 montyred(u64 lo, u64 hi) {
    u32 lo0 = (u32)lo, lo1 = (u32)(lo >> 32);
    u32 e, k0, k1, c0, c1, d0, d1;
    u32 a1 = __builtin_addc(lo1, lo0, 0u, &e);
    u32 b0 = __builtin_subc(lo0, a1, e, &k0);
    u32 b1 = __builtin_subc(a1, 0u, k0, &k1);
    u32 r0 = __builtin_subc((u32)hi, b0, 0u, &c0);
    u32 r1 = __builtin_subc((u32)(hi >> 32), b1, c0, &c1);
    r0 = __builtin_subc(r0, 0u - c1, 0u, &d0);
    r1 = __builtin_subc(r1, 0u, d0, &d1);
    return mk64(r0, r1);
}

// L + H * 2^32 mod p for L, H < 2^52 (the MDS recombination) -> < 2^64, in
// 32-bit words: with 2^64 = 2^32 - 1 mod p the value is (L0 - H1) +
// (L1 + H0 + H1) * 2^32; a carry out of the high word is folded back as 2^32 - 1.
__host__ __device__ __forceinline__ u64 lin52(u64 L, u64 H) {
    u32 L0 = (u32)L, L1 = (u32)(L >> 32), H0 = (u32)H, H1 = (u32)(H >> 32);
    u32 b, b2, c, c2, d, d2;
    u32 lo = __builtin_subc(L0, H1, 0u, &b);
    u32 t = __builtin_subc(L1 + H1, 0u, b, &b2);  // L1 + H1 - b >= 0: b = 1 needs H1 > 0
    u32 h = __builtin_addc(H0, t, 0u, &c);       // high word and its carry (value 2^64)
    // carry: add 2^64 = 2^32 - 1 (mod p); then h < 2^21, so h + 1 cannot overflow
    lo = __builtin_subc(lo, 0u, c, &d);
    h = __builtin_addc(h, 0u, c, &c2);
    h = __builtin_subc(h, 0u, d, &d2);
    (void)b2;
    (void)c2;
    (void)d2;
    return mk64(lo, h);
}

// a * a * 2^-64 mod p, any u64 a -> < 2^64
__host__ __device__ __forceinline__ u64 mont_sqr(u64 a) {
    u64 lo, hi;
    sqr_wide(a, lo, hi);
    return montyred(lo, hi);
}

// x * 2^64 mod p (into Montgomery form), any u64 x -> < 2^64:
// x*2^64 = x*(2^32 - 1) = x0*2^32 - x0 - x1 (mod p) for x = x1*2^32 + x0.
__host__ __device__ __forceinline__ u64 to_mont(u64 x) {
    u32 x0 = (u32)x, x1 = (u32)(x >> 32);
    u64 t = (u64)x0 * EPS;  // < 2^64
    u64 r = t - x1;
    if (t < x1) r -= EPS;   // borrow (only when x0 = 0): add p
    return r;
}

// a * b * 2^-64 mod p (Montgomery product), any u64 inputs -> < 2^64
__host__ __device__ __forceinline__ u64 mont_mul(u64 a, u64 b) {
    u64 lo, hi;
    mul_wide(a, b, lo, hi);
    return montyred(lo, hi);
}

// a, b any u64 -> lazy
__host__ __device__ __forceinline__ u64 mul_lazy(u64 a, u64 b) {
    u64 lo, hi;
    mul_wide(a, b, lo, hi);
    return fold128(lo, hi);
}

// a, b any u64 -> canonical
__host__ __device__ __forceinline__ u64 mul(u64 a, u64 b) { return canon_raw(mul_lazy(a, b)); }

// 2^e mod p for 0 <= e < 192 (compile-time usable)
__host__ __device__ constexpr u64 pow2_mod_p(int e) {
    u64 r = 1;
    for (int i = 0; i < e; i++) r = (r >= P - r) ? r - (P - r) : r + r;
    return r;
}

// x any u64, 0 < E <= 32 -> lazy x * 2^E: lo = x << E, hl = x >> (64 - E) < 2^32
template <int E>
__host__ __device__ __forceinline__ u64 mul_pow2_small(u64 x) {
    static_assert(E > 0 && E <= 32, "small shift");
    u64 lo = x << E;
    u32 hl = (u32)(x >> (64 - E));
    u32 k0, k1, c0, c1, c2, c3;
    u32 u0 = __builtin_subc(0u, hl, 0u, &k0);
    u32 u1 = __builtin_subc(hl, 0u, k0, &k1);
    u32 r0 = __builtin_addc((u32)lo, u0, 0u, &c0);
    u32 r1 = __builtin_addc((u32)(lo >> 32), u1, c0, &c1);
    r0 = __builtin_addc(r0, c1 ? 0xFFFFFFFFu : 0u, 0u, &c2);
    r1 = __builtin_addc(r1, 0u, c2, &c3);
    return mk64(r0, r1);
}

// x any u64 -> lazy x * 2^32
__host__ __device__ __forceinline__ u64 mul_2_32(u64 x) {
    // x = x1*2^32 + x0 ; x*2^32 = x0*2^32 + x1*2^64 = x0*2^32 + x1*(2^32-1)
    return fold128(x << 32, x >> 32);
}

// x * 2^E mod p for a compile-time 0 <= E < 96 and any u64 x -> canonical,
// without a general product.  p divides 2^96 + 1, so with x << E split into
// 32-bit words w0..w4 (E = 32q + r), the 96 low bits count positively and the
// words above bit 96 negatively: x*2^E = lo + w2*2^64 - hi, lo = (w1:w0),
// hi = (w4:w3), and 2^64 = 2^32 - 1 mod p turns w2*2^64 into one
// multiply-add.  The word shuffles are free; r != 0 costs funnel shifts.
template <int E>
__host__ __device__ __forceinline__ u64 mul_pow2(u64 x) {
    static_assert(E >= 0 && E < 96, "exponent");
    constexpr int q = E / 32, r = E % 32;
    const u32 x0 = (u32)x, x1 = (u32)(x >> 32);
    u32 w[5] = {0, 0, 0, 0, 0};
    if (r == 0) {
        w[q] = x0;
        w[q + 1] = x1;
    } else {
        w[q] = x0 << r;
        w[q + 1] = (x1 << r) | (x0 >> (32 - r));
        w[q + 2] = x1 >> (32 - r);
    }
    // In 32-bit words (no 64-bit multiply or compare):
    // m = w2 * (2^32 - 1) = (w2 - [w2 != 0]) : (-w2)
    u32 b0, b1, c0, c1, c2, c3;
    u32 m0 = __builtin_subc(0u, w[2], 0u, &b0);
    u32 m1 = __builtin_subc(w[2], 0u, b0, &b1);
    // t = lo + m < 2^65 - 2^33; a carry (2^64) folds back as EPS and cannot
    // carry again (t mod 2^64 <= 2^64 - 2^33 then)
    u32 t0 = __builtin_addc(w[0], m0, 0u, &c0);
    u32 t1 = __builtin_addc(w[1], m1, c0, &c1);
    t0 = __builtin_addc(t0, c1 ? 0xFFFFFFFFu : 0u, 0u, &c2);
    t1 = __builtin_addc(t1, 0u, c2, &c3);
    // v = t - hi, hi = (w4:w3) = x * 2^E >> 96 < 2^63 (0 when E < 32), so
    // -2^63 < v < 2^64; the canonical result is one of u = v mod 2^64, u + EPS
    // (v >= p: minus p) and u - EPS (v < 0: plus p, which lands in (0, p)),
    // selected at the end (no borrow fix feeding a second carry chain: the
    // gfx950 compiler merges such chains incorrectly, see opaque())
    u32 d1 = 0, u0 = t0, u1 = t1;
    if (q > 0) {
        u32 d0;
        u0 = __builtin_subc(t0, w[3], 0u, &d0);
        u1 = __builtin_subc(t1, w[4], d0, &d1);
    }
    u32 k0, k1, g0, g1;
    u32 a0 = __builtin_addc(u0, 0xFFFFFFFFu, 0u, &k0);  // u + EPS carries iff u >= p
    u32 a1 = __builtin_addc(u1, 0u, k0, &k1);
    u32 s0 = __builtin_subc(u0, 0xFFFFFFFFu, 0u, &g0);  // u - EPS (mod 2^64)
    u32 s1 = __builtin_subc(u1, 0u, g0, &g1);
    (void)b1;
    (void)c3;
    (void)g1;
    return d1 ? mk64(s0, s1) : (k1 ? mk64(a0, a1) : mk64(u0, u1));
}

// x canonical, k runtime in [0, 192) -> canonical x * 2^k (host and tables)
__host__ __device__ __forceinline__ u64 mul_pow2_rt(u64 x, u32 k) {
    bool negate = k >= 96;
    if (negate) k -= 96;
    u64 r;
    if (k == 0) {
        r = x;
    } else if (k < 64) {
        r = canon(fold128(x << k, x >> (64 - k)));
    } else {
        u32 s = k - 64;
        u64 y = s ? canon(fold128(x << s, x >> (64 - s))) : x;
        r = sub(canon(mul_2_32(y)), y);
    }
    return negate ? neg(r) : r;
}

__host__ __device__ inline u64 pow(u64 b, u64 e) {
    u64 r = 1;
    while (e) {
        if (e & 1) r = mul(r, b);
        b = mul(b, b);
        e >>= 1;
    }
    return r;
}

__host__ __device__ inline u64 inv(u64 a) { return pow(a, P - 2); }

// The reference's primitive 2^k-th root of unity, 7^((p-1)/2^k)
// (src/types/b_field_element.cpp:179-192).
__host__ __device__ inline u64 root_of_unity(u32 k) { return pow(7, (P - 1) >> k); }

}  // namespace gl

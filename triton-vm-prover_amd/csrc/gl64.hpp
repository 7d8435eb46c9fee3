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

__host__ __device__ __forceinline__ u64 canon(u64 x) { return x >= P ? x - P : x; }

// a, b canonical -> canonical
__host__ __device__ __forceinline__ u64 add(u64 a, u64 b) {
    u64 s = a + b;
    // carry (s < a) or s >= p: subtract p, i.e. add EPS modulo 2^64
    return (s < a || s >= P) ? s + EPS : s;
}

// a, b canonical -> canonical
__host__ __device__ __forceinline__ u64 sub(u64 a, u64 b) {
    u64 d = a - b;
    return (a < b) ? d - EPS : d;
}

// a canonical -> canonical
__host__ __device__ __forceinline__ u64 neg(u64 a) { return a ? P - a : 0; }

// a any u64, b canonical -> lazy (any u64).  Used to add round constants.
__host__ __device__ __forceinline__ u64 add_lazy(u64 a, u64 b) {
    u64 s = a + b;
    return (s < a) ? s + EPS : s;  // a + b < 2^65 - 2^32: one fold suffices
}

// 128-bit value (hi:lo) -> lazy u64 congruent to it.
__host__ __device__ __forceinline__ u64 fold128(u64 lo, u64 hi) {
    u32 hh = (u32)(hi >> 32), hl = (u32)hi;
    u64 t = lo - hh;                       // + hh * 2^96 = - hh
    if (lo < hh) t -= EPS;                 // borrow: add p
    u64 u = ((u64)hl << 32) - hl;          // hl * 2^64 = hl * (2^32 - 1)
    u64 r = t + u;
    if (r < u) r += EPS;                   // carry: subtract p
    return r;
}

__host__ __device__ __forceinline__ void mul_wide(u64 a, u64 b, u64& lo, u64& hi) {
#ifdef __HIP_DEVICE_COMPILE__
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 p00 = (u64)a0 * b0;
    u64 m1 = (u64)a0 * b1 + (p00 >> 32);   // < 2^64
    u64 m2 = (u64)a1 * b0 + (u32)m1;       // < 2^64
    lo = (m2 << 32) | (u32)p00;
    hi = (u64)a1 * b1 + (m1 >> 32) + (m2 >> 32);
#else
    unsigned __int128 pr = (unsigned __int128)a * b;
    lo = (u64)pr;
    hi = (u64)(pr >> 64);
#endif
}

// a, b any u64 -> lazy
__host__ __device__ __forceinline__ u64 mul_lazy(u64 a, u64 b) {
    u64 lo, hi;
    mul_wide(a, b, lo, hi);
    return fold128(lo, hi);
}

// a, b any u64 -> canonical
__host__ __device__ __forceinline__ u64 mul(u64 a, u64 b) { return canon(mul_lazy(a, b)); }

// x any u64 -> lazy x * 2^32
__host__ __device__ __forceinline__ u64 mul_2_32(u64 x) {
    // x = x1*2^32 + x0 ; x*2^32 = x0*2^32 + x1*2^64 = x0*2^32 + x1*(2^32-1)
    return fold128(x << 32, x >> 32);
}

// x canonical, K compile-time in [0, 192) -> canonical x * 2^K
template <int K>
__host__ __device__ __forceinline__ u64 mul_pow2(u64 x) {
    static_assert(K >= 0 && K < 192, "shift out of range");
    if constexpr (K == 0) {
        return x;
    } else if constexpr (K >= 96) {
        return neg(mul_pow2<K - 96>(x));   // 2^96 = -1
    } else if constexpr (K < 64) {
        return canon(fold128(x << K, x >> (64 - K)));
    } else {
        // 2^K = 2^(K-64) * 2^64 ; 2^64 = 2^32 - 1
        u64 y = (K == 64) ? x : canon(fold128(x << (K - 64), x >> (128 - K)));
        return sub(canon(mul_2_32(y)), y);
    }
}

// x canonical, k runtime in [0, 192) -> canonical x * 2^k
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

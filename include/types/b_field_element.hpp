// BFieldElement: an element of the Goldilocks field p = 2^64 - 2^32 + 1,
// stored as its canonical u64 in [0, p).
//
// Drop-in for the prover's value type (include/types/b_field_element.hpp:28-104,
// src/types/b_field_element.cpp): same class, members, signatures, exceptions
// and memory layout (one u64), so code written against that header compiles
// unchanged with this one first on the include path.  Header-only: the bulk
// work that uses these values runs on the GPU (ntt/ntt.hpp, lde/*.hpp,
// hash/tip5.hpp, merkle/merkle_tree.hpp); single-element arithmetic is plain
// host arithmetic of the value type.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace triton_vm {

using uint128_t = __uint128_t;

class BFieldElement {
public:
    static constexpr uint64_t MODULUS = 18446744069414584321ULL;
    // 2^64 mod p: the Montgomery radix R reduced, R = 2^64
    static constexpr uint64_t MONTGOMERY_R = 4294967295ULL;
    static constexpr uint64_t GENERATOR = 7ULL;

    constexpr BFieldElement() : value_(0) {}
    // reduces any u64 modulo p (b_field_element.hpp:41)
    constexpr explicit BFieldElement(uint64_t value) : value_(value % MODULUS) {}
    // the element whose Montgomery form (x * 2^64 mod p) is montgomery_value
    static BFieldElement from_raw_u64(uint64_t montgomery_value) {
        return BFieldElement(montyred(static_cast<uint128_t>(montgomery_value)));
    }

    static constexpr BFieldElement zero() { return BFieldElement(0); }
    static constexpr BFieldElement one() { return BFieldElement(1); }
    static constexpr BFieldElement generator() { return BFieldElement(GENERATOR); }

    constexpr uint64_t value() const { return value_; }

    BFieldElement operator+(const BFieldElement& rhs) const {
        uint64_t s = value_ + rhs.value_;
        // a carry out of 64 bits, or s >= p, means subtract p once
        bool over = s < value_ || s >= MODULUS;
        return raw(over ? s - MODULUS : s);
    }
    BFieldElement operator-(const BFieldElement& rhs) const {
        return raw(value_ >= rhs.value_ ? value_ - rhs.value_ : value_ + (MODULUS - rhs.value_));
    }
    BFieldElement operator*(const BFieldElement& rhs) const {
        return raw(reduce(static_cast<uint128_t>(value_) * rhs.value_));
    }
    BFieldElement operator/(const BFieldElement& rhs) const { return *this * rhs.inverse(); }
    BFieldElement operator-() const { return raw(value_ ? MODULUS - value_ : 0); }

    BFieldElement& operator+=(const BFieldElement& rhs) { return *this = *this + rhs; }
    BFieldElement& operator-=(const BFieldElement& rhs) { return *this = *this - rhs; }
    BFieldElement& operator*=(const BFieldElement& rhs) { return *this = *this * rhs; }
    BFieldElement& operator/=(const BFieldElement& rhs) { return *this = *this / rhs; }

    bool operator==(const BFieldElement& rhs) const { return value_ == rhs.value_; }
    bool operator!=(const BFieldElement& rhs) const { return value_ != rhs.value_; }
    bool operator<(const BFieldElement& rhs) const { return value_ < rhs.value_; }
    bool operator>(const BFieldElement& rhs) const { return value_ > rhs.value_; }
    bool operator<=(const BFieldElement& rhs) const { return value_ <= rhs.value_; }
    bool operator>=(const BFieldElement& rhs) const { return value_ >= rhs.value_; }

    // throws std::domain_error for zero (b_field_element.cpp)
    BFieldElement inverse() const {
        if (value_ == 0) throw std::domain_error("Cannot invert zero");
        return pow(MODULUS - 2);
    }
    BFieldElement pow(uint64_t exp) const {
        BFieldElement r = one(), b = *this;
        for (; exp; exp >>= 1) {
            if (exp & 1) r *= b;
            b *= b;
        }
        return r;
    }
    bool is_zero() const { return value_ == 0; }
    bool is_one() const { return value_ == 1; }

    // Montgomery's trick: one inversion and 3(n-1) products; throws
    // std::domain_error when an element is zero
    static std::vector<BFieldElement> batch_inversion(const std::vector<BFieldElement>& elements) {
        const size_t n = elements.size();
        std::vector<BFieldElement> out(n);
        if (n == 0) return out;
        BFieldElement acc = one();
        for (size_t i = 0; i < n; i++) {
            if (elements[i].is_zero()) throw std::domain_error("batch_inversion encountered zero element");
            out[i] = acc;  // product of elements[0..i)
            acc *= elements[i];
        }
        BFieldElement inv = acc.inverse();
        for (size_t i = n; i-- > 0;) {
            out[i] *= inv;
            inv *= elements[i];
        }
        return out;
    }

    static void batch_add(BFieldElement* dst, const BFieldElement* a, const BFieldElement* b,
                          size_t count) {
        for (size_t i = 0; i < count; i++) dst[i] = a[i] + b[i];
    }
    static void batch_mul(BFieldElement* dst, const BFieldElement* a, const BFieldElement* b,
                          size_t count) {
        for (size_t i = 0; i < count; i++) dst[i] = a[i] * b[i];
    }
    static void batch_sub(BFieldElement* dst, const BFieldElement* a, const BFieldElement* b,
                          size_t count) {
        for (size_t i = 0; i < count; i++) dst[i] = a[i] - b[i];
    }

    // 7^((p-1)/2^k), k <= 32 (b_field_element.cpp:179-192)
    static BFieldElement primitive_root_of_unity(uint32_t log2_order) {
        if (log2_order > 32) throw std::invalid_argument("log2_order must be <= 32");
        return BFieldElement(GENERATOR).pow((MODULUS - 1) >> log2_order);
    }

    std::string to_string() const { return std::to_string(value_); }
    friend std::ostream& operator<<(std::ostream& os, const BFieldElement& elem) {
        return os << elem.value_;
    }

private:
    uint64_t value_;

    static BFieldElement raw(uint64_t canonical) {
        BFieldElement e;
        e.value_ = canonical;
        return e;
    }
    // 128-bit value mod p: lo + hl*2^64 + hh*2^96 = lo + hl*(2^32 - 1) - hh
    static uint64_t reduce(uint128_t v) {
        uint64_t lo = static_cast<uint64_t>(v);
        uint64_t hi = static_cast<uint64_t>(v >> 64);
        uint64_t hh = hi >> 32, hl = hi & 0xFFFFFFFFull;
        uint64_t t = lo - hh;
        if (lo < hh) t -= 0xFFFFFFFFull;  // borrow: add p (mod 2^64)
        uint64_t u = (hl << 32) - hl;
        uint64_t r = t + u;
        if (r < u) r += 0xFFFFFFFFull;  // carry: subtract p (mod 2^64)
        return r >= MODULUS ? r - MODULUS : r;
    }
    // v * 2^-64 mod p for v < 2^128 with v / 2^64 < p (Montgomery reduction);
    // canonical for such inputs
    static uint64_t montyred(uint128_t v) {
        uint64_t lo = static_cast<uint64_t>(v), hi = static_cast<uint64_t>(v >> 64);
        uint64_t m = lo + (lo << 32);         // lo * (1 + 2^32) = lo * (-p^-1) mod 2^64
        bool over = m < lo;                   // the carry of that sum (value 2^64)
        uint64_t q = m - (m >> 32) - over;    // (m * p + lo) / 2^64 - (the 2^64 carry)
        uint64_t r = hi - q;
        return hi < q ? r + MODULUS : r;
    }
};

using BFE = BFieldElement;

}  // namespace triton_vm

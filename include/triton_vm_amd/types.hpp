// Value types of the drop-in C++ interface (namespace triton_vm).
//
// Same names, layouts and semantics as the reference prover's types so code
// written against them compiles unchanged:
//   BFieldElement   include/types/b_field_element.hpp:28-104 (canonical u64, ctor reduces mod p)
//   XFieldElement   include/types/x_field_element.hpp:18-86  (c0 + c1 X + c2 X^2, X^3 = X - 1)
//   Digest          include/types/digest.hpp:16-55            (5 elements, LE hex)
//   ArithmeticDomain include/table/master_table.hpp:49-59     (length, offset, generator)
// Arithmetic on single elements is host arithmetic of the value type; bulk
// work (NTT, LDE, hashing, Merkle) goes through the GPU library (ntt.hpp,
// lde.hpp, tip5.hpp, merkle.hpp).
#pragma once

#include <array>
#include <cstdint>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace triton_vm {

class BFieldElement {
public:
    static constexpr uint64_t MODULUS = 18446744069414584321ULL;
    static constexpr uint64_t GENERATOR = 7ULL;

    constexpr BFieldElement() : value_(0) {}
    constexpr explicit BFieldElement(uint64_t v) : value_(v % MODULUS) {}

    static constexpr BFieldElement zero() { return BFieldElement(0); }
    static constexpr BFieldElement one() { return BFieldElement(1); }
    static constexpr BFieldElement generator() { return BFieldElement(GENERATOR); }

    constexpr uint64_t value() const { return value_; }
    bool is_zero() const { return value_ == 0; }
    bool is_one() const { return value_ == 1; }

    BFieldElement operator+(const BFieldElement& r) const {
        uint64_t s = value_ + r.value_;
        return raw((s < value_ || s >= MODULUS) ? s - MODULUS : s);
    }
    BFieldElement operator-(const BFieldElement& r) const {
        return raw(value_ >= r.value_ ? value_ - r.value_ : value_ + (MODULUS - r.value_));
    }
    BFieldElement operator*(const BFieldElement& r) const {
        return raw((uint64_t)(((unsigned __int128)value_ * r.value_) % MODULUS));
    }
    BFieldElement operator-() const { return raw(value_ ? MODULUS - value_ : 0); }
    BFieldElement operator/(const BFieldElement& r) const { return *this * r.inverse(); }
    BFieldElement& operator+=(const BFieldElement& r) { return *this = *this + r; }
    BFieldElement& operator-=(const BFieldElement& r) { return *this = *this - r; }
    BFieldElement& operator*=(const BFieldElement& r) { return *this = *this * r; }
    BFieldElement& operator/=(const BFieldElement& r) { return *this = *this / r; }
    bool operator==(const BFieldElement& r) const { return value_ == r.value_; }
    bool operator!=(const BFieldElement& r) const { return value_ != r.value_; }
    bool operator<(const BFieldElement& r) const { return value_ < r.value_; }

    BFieldElement pow(uint64_t e) const {
        BFieldElement r = one(), b = *this;
        while (e) {
            if (e & 1) r *= b;
            b *= b;
            e >>= 1;
        }
        return r;
    }
    BFieldElement inverse() const {
        if (!value_) throw std::domain_error("Cannot invert zero");
        return pow(MODULUS - 2);
    }
    // 7^((p-1)/2^k) (src/types/b_field_element.cpp:179-192)
    static BFieldElement primitive_root_of_unity(uint32_t log2_order) {
        if (log2_order > 32) throw std::invalid_argument("log2_order must be <= 32");
        return BFieldElement(GENERATOR).pow((MODULUS - 1) >> log2_order);
    }
    std::string to_string() const { return std::to_string(value_); }
    friend std::ostream& operator<<(std::ostream& os, const BFieldElement& e) {
        return os << e.value_;
    }

private:
    static BFieldElement raw(uint64_t v) {
        BFieldElement e;
        e.value_ = v;
        return e;
    }
    uint64_t value_;
};

using BFE = BFieldElement;

class XFieldElement {
public:
    constexpr XFieldElement() : coeffs_{} {}
    constexpr XFieldElement(BFieldElement c0, BFieldElement c1, BFieldElement c2)
        : coeffs_{c0, c1, c2} {}
    constexpr explicit XFieldElement(BFieldElement b) : coeffs_{b, BFieldElement(), BFieldElement()} {}

    const std::array<BFieldElement, 3>& coefficients() const { return coeffs_; }
    BFieldElement coeff(size_t i) const { return coeffs_[i]; }

    XFieldElement operator+(const XFieldElement& r) const {
        return {coeffs_[0] + r.coeffs_[0], coeffs_[1] + r.coeffs_[1], coeffs_[2] + r.coeffs_[2]};
    }
    XFieldElement operator-(const XFieldElement& r) const {
        return {coeffs_[0] - r.coeffs_[0], coeffs_[1] - r.coeffs_[1], coeffs_[2] - r.coeffs_[2]};
    }
    XFieldElement operator*(const BFieldElement& s) const {
        return {coeffs_[0] * s, coeffs_[1] * s, coeffs_[2] * s};
    }
    // (a0 + a1 X + a2 X^2)(b0 + b1 X + b2 X^2) mod X^3 - X + 1 (x_field_element.cpp:24-48)
    XFieldElement operator*(const XFieldElement& r) const {
        const auto& a = coeffs_;
        const auto& b = r.coeffs_;
        BFieldElement c0 = a[0] * b[0];
        BFieldElement c1 = a[0] * b[1] + a[1] * b[0];
        BFieldElement c2 = a[0] * b[2] + a[1] * b[1] + a[2] * b[0];
        BFieldElement c3 = a[1] * b[2] + a[2] * b[1];
        BFieldElement c4 = a[2] * b[2];
        return {c0 - c3, c1 + c3 - c4, c2 + c4};
    }
    bool operator==(const XFieldElement& r) const { return coeffs_ == r.coeffs_; }
    bool operator!=(const XFieldElement& r) const { return !(*this == r); }

private:
    std::array<BFieldElement, 3> coeffs_;
};

class Digest {
public:
    static constexpr size_t LEN = 5;
    Digest() : elements_{} {}
    explicit Digest(const std::array<BFieldElement, LEN>& e) : elements_(e) {}
    Digest(BFieldElement e0, BFieldElement e1, BFieldElement e2, BFieldElement e3, BFieldElement e4)
        : elements_{e0, e1, e2, e3, e4} {}
    static Digest zero() { return Digest(); }
    const std::array<BFieldElement, LEN>& elements() const { return elements_; }
    BFieldElement operator[](size_t i) const { return elements_[i]; }
    BFieldElement& operator[](size_t i) { return elements_[i]; }
    std::vector<BFieldElement> to_b_field_elements() const {
        return std::vector<BFieldElement>(elements_.begin(), elements_.end());
    }
    bool operator==(const Digest& r) const { return elements_ == r.elements_; }
    bool operator!=(const Digest& r) const { return !(*this == r); }
    // 8 little-endian bytes per element (src/types/digest.cpp:21-61)
    std::string to_hex() const {
        static const char* hexd = "0123456789abcdef";
        std::string s;
        for (const auto& e : elements_)
            for (int b = 0; b < 8; b++) {
                unsigned v = (unsigned)((e.value() >> (8 * b)) & 0xff);
                s += hexd[v >> 4];
                s += hexd[v & 15];
            }
        return s;
    }
    static Digest from_hex(const std::string& hex) {
        if (hex.size() != LEN * 16) throw std::invalid_argument("Invalid hex string length for Digest");
        std::array<BFieldElement, LEN> e;
        for (size_t i = 0; i < LEN; i++) {
            uint64_t v = 0;
            for (int b = 0; b < 8; b++)
                v |= (uint64_t)std::stoul(hex.substr(16 * i + 2 * b, 2), nullptr, 16) << (8 * b);
            e[i] = BFieldElement(v);
        }
        return Digest(e);
    }
    friend std::ostream& operator<<(std::ostream& os, const Digest& d) { return os << d.to_hex(); }

private:
    std::array<BFieldElement, LEN> elements_;
};

struct ArithmeticDomain {
    size_t length = 0;
    BFieldElement offset = BFieldElement::one();
    BFieldElement generator = BFieldElement::one();

    // (length, offset 1, w_length) (src/table/master_table.cpp:309-323)
    static ArithmeticDomain of_length(size_t length) {
        if (length == 0 || (length & (length - 1)))
            throw std::invalid_argument("domain length must be a power of 2");
        uint32_t lg = 0;
        while ((size_t(1) << lg) < length) lg++;
        ArithmeticDomain d;
        d.length = length;
        d.offset = BFieldElement::one();
        d.generator = BFieldElement::primitive_root_of_unity(lg);
        return d;
    }
    ArithmeticDomain with_offset(BFieldElement o) const {
        ArithmeticDomain d = *this;
        d.offset = o;
        return d;
    }
    BFieldElement element(size_t i) const { return offset * generator.pow(i); }
    std::vector<BFieldElement> values() const {
        std::vector<BFieldElement> v(length);
        BFieldElement x = offset;
        for (size_t i = 0; i < length; i++, x *= generator) v[i] = x;
        return v;
    }
};

}  // namespace triton_vm

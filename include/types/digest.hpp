// Digest: five BFieldElements, the output of Tip5 hashing.
//
// Drop-in for the prover's digest type (include/types/digest.hpp:16-55,
// src/types/digest.cpp:21-61): same members and layout (five u64); the hex
// form writes each element as 8 little-endian bytes.
#pragma once

#include <array>
#include <cstddef>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "types/b_field_element.hpp"

namespace triton_vm {

class Digest {
public:
    static constexpr size_t LEN = 5;

    Digest() : elements_{} {}
    explicit Digest(const std::array<BFieldElement, LEN>& elements) : elements_(elements) {}
    Digest(BFieldElement e0, BFieldElement e1, BFieldElement e2, BFieldElement e3,
           BFieldElement e4)
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

    std::string to_hex() const {
        static const char kHex[] = "0123456789abcdef";
        std::string s;
        s.reserve(LEN * 16);
        for (const BFieldElement& e : elements_) {
            uint64_t v = e.value();
            for (int byte = 0; byte < 8; byte++, v >>= 8) {
                s += kHex[(v >> 4) & 15];
                s += kHex[v & 15];
            }
        }
        return s;
    }
    // inverse of to_hex; throws std::invalid_argument for a wrong length or a
    // non-hex character
    static Digest from_hex(const std::string& hex) {
        if (hex.size() != LEN * 16) throw std::invalid_argument("Invalid hex string length for Digest");
        auto nibble = [](char c) -> uint64_t {
            if (c >= '0' && c <= '9') return (uint64_t)(c - '0');
            if (c >= 'a' && c <= 'f') return (uint64_t)(c - 'a' + 10);
            if (c >= 'A' && c <= 'F') return (uint64_t)(c - 'A' + 10);
            throw std::invalid_argument("Invalid hex character in Digest");
        };
        std::array<BFieldElement, LEN> e;
        for (size_t i = 0; i < LEN; i++) {
            uint64_t v = 0;
            for (int byte = 0; byte < 8; byte++) {
                size_t at = 16 * i + 2 * byte;
                v |= ((nibble(hex[at]) << 4) | nibble(hex[at + 1])) << (8 * byte);
            }
            e[i] = BFieldElement(v);
        }
        return Digest(e);
    }

    friend std::ostream& operator<<(std::ostream& os, const Digest& d) { return os << d.to_hex(); }

private:
    std::array<BFieldElement, LEN> elements_;
};

}  // namespace triton_vm

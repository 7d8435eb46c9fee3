// XFieldElement: c0 + c1 X + c2 X^2 in F_p[X] / (X^3 - X + 1).
//
// Drop-in for the prover's extension-field value type
// (include/types/x_field_element.hpp:18-86, src/types/x_field_element.cpp):
// same members, signatures, exceptions, string form "(c0, c1, c2)" and
// layout (three BFieldElements).  Header-only host arithmetic.
#pragma once

#include <array>
#include <cstddef>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "types/b_field_element.hpp"

namespace triton_vm {

class XFieldElement {
public:
    static constexpr size_t EXTENSION_DEGREE = 3;

    constexpr XFieldElement()
        : coeffs_{BFieldElement::zero(), BFieldElement::zero(), BFieldElement::zero()} {}
    constexpr XFieldElement(BFieldElement c0, BFieldElement c1, BFieldElement c2)
        : coeffs_{c0, c1, c2} {}
    constexpr explicit XFieldElement(BFieldElement base)
        : coeffs_{base, BFieldElement::zero(), BFieldElement::zero()} {}

    static constexpr XFieldElement zero() { return XFieldElement(); }
    static constexpr XFieldElement one() {
        return XFieldElement(BFieldElement::one(), BFieldElement::zero(), BFieldElement::zero());
    }

    constexpr const std::array<BFieldElement, 3>& coefficients() const { return coeffs_; }
    constexpr BFieldElement coeff(size_t i) const { return coeffs_[i]; }

    XFieldElement operator+(const XFieldElement& r) const {
        return {coeffs_[0] + r.coeffs_[0], coeffs_[1] + r.coeffs_[1], coeffs_[2] + r.coeffs_[2]};
    }
    XFieldElement operator-(const XFieldElement& r) const {
        return {coeffs_[0] - r.coeffs_[0], coeffs_[1] - r.coeffs_[1], coeffs_[2] - r.coeffs_[2]};
    }
    // schoolbook product, then X^3 = X - 1 and X^4 = X^2 - X
    XFieldElement operator*(const XFieldElement& r) const {
        const auto& a = coeffs_;
        const auto& b = r.coeffs_;
        BFieldElement d0 = a[0] * b[0];
        BFieldElement d1 = a[0] * b[1] + a[1] * b[0];
        BFieldElement d2 = a[0] * b[2] + a[1] * b[1] + a[2] * b[0];
        BFieldElement d3 = a[1] * b[2] + a[2] * b[1];
        BFieldElement d4 = a[2] * b[2];
        return {d0 - d3, d1 + d3 - d4, d2 + d4};
    }
    XFieldElement operator/(const XFieldElement& r) const { return *this * r.inverse(); }
    XFieldElement operator-() const { return {-coeffs_[0], -coeffs_[1], -coeffs_[2]}; }

    XFieldElement& operator+=(const XFieldElement& r) { return *this = *this + r; }
    XFieldElement& operator-=(const XFieldElement& r) { return *this = *this - r; }
    XFieldElement& operator*=(const XFieldElement& r) { return *this = *this * r; }
    XFieldElement& operator/=(const XFieldElement& r) { return *this = *this / r; }

    XFieldElement operator+(const BFieldElement& r) const {
        return {coeffs_[0] + r, coeffs_[1], coeffs_[2]};
    }
    XFieldElement operator-(const BFieldElement& r) const {
        return {coeffs_[0] - r, coeffs_[1], coeffs_[2]};
    }
    XFieldElement operator*(const BFieldElement& r) const {
        return {coeffs_[0] * r, coeffs_[1] * r, coeffs_[2] * r};
    }
    XFieldElement operator/(const BFieldElement& r) const { return *this * r.inverse(); }

    bool operator==(const XFieldElement& r) const { return coeffs_ == r.coeffs_; }
    bool operator!=(const XFieldElement& r) const { return !(*this == r); }

    // Multiplication by a is the linear map with matrix (columns: images of
    // 1, X, X^2)
    //   | a0   -a2      -a1     |
    //   | a1   a0 + a2  a1 - a2 |
    //   | a2   a1       a0 + a2 |
    // so a^-1 = M^-1 e0 = (first column of adj(M)) / det(M).  Throws
    // std::domain_error for zero.
    XFieldElement inverse() const {
        if (is_zero()) throw std::domain_error("Cannot invert zero");
        const BFieldElement a0 = coeffs_[0], a1 = coeffs_[1], a2 = coeffs_[2];
        const BFieldElement m00 = a0, m01 = -a2, m02 = -a1;
        const BFieldElement m10 = a1, m11 = a0 + a2, m12 = a1 - a2;
        const BFieldElement m20 = a2, m21 = a1, m22 = a0 + a2;
        // cofactors of the first row = first column of the adjugate
        const BFieldElement c00 = m11 * m22 - m12 * m21;
        const BFieldElement c01 = m12 * m20 - m10 * m22;
        const BFieldElement c02 = m10 * m21 - m11 * m20;
        const BFieldElement det = m00 * c00 + m01 * c01 + m02 * c02;
        const BFieldElement dinv = det.inverse();
        return {c00 * dinv, c01 * dinv, c02 * dinv};
    }
    XFieldElement pow(uint64_t exp) const {
        XFieldElement r = one(), b = *this;
        for (; exp; exp >>= 1) {
            if (exp & 1) r *= b;
            b *= b;
        }
        return r;
    }
    bool is_zero() const { return coeffs_[0].is_zero() && coeffs_[1].is_zero() && coeffs_[2].is_zero(); }
    bool is_one() const { return coeffs_[0].is_one() && coeffs_[1].is_zero() && coeffs_[2].is_zero(); }

    std::string to_string() const {
        return "(" + std::to_string(coeffs_[0].value()) + ", " + std::to_string(coeffs_[1].value()) +
               ", " + std::to_string(coeffs_[2].value()) + ")";
    }

    // Montgomery's trick over the extension field; throws std::domain_error
    // when an element is zero
    static std::vector<XFieldElement> batch_inversion(const std::vector<XFieldElement>& elements) {
        const size_t n = elements.size();
        std::vector<XFieldElement> out(n);
        if (n == 0) return out;
        XFieldElement acc = one();
        for (size_t i = 0; i < n; i++) {
            if (elements[i].is_zero()) throw std::domain_error("Cannot invert zero");
            out[i] = acc;
            acc *= elements[i];
        }
        XFieldElement inv = acc.inverse();
        for (size_t i = n; i-- > 0;) {
            out[i] *= inv;
            inv *= elements[i];
        }
        return out;
    }

    friend std::ostream& operator<<(std::ostream& os, const XFieldElement& e) {
        return os << e.to_string();
    }
    friend XFieldElement operator+(const BFieldElement& l, const XFieldElement& r) { return r + l; }
    friend XFieldElement operator-(const BFieldElement& l, const XFieldElement& r) {
        return XFieldElement(l) - r;
    }
    friend XFieldElement operator*(const BFieldElement& l, const XFieldElement& r) { return r * l; }

private:
    std::array<BFieldElement, 3> coeffs_;
};

using XFE = XFieldElement;

}  // namespace triton_vm

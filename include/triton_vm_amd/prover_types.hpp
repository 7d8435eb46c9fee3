// ArithmeticDomain and Polynomial for builds without the prover's own headers.
//
// The drop-in headers (ntt/ntt.hpp, lde/*.hpp, table/table_commitment.hpp)
// take these two types from the prover's table/master_table.hpp and
// polynomial/polynomial.hpp when those headers are on the include path (a
// prover build with this directory first).  Standalone builds (this library,
// its tests) get the definitions below instead.  The layouts are the same as
// the prover's (master_table.hpp:49-59: length, offset, generator;
// polynomial.hpp: one coefficient vector, lowest degree first), which is what
// crosses the library boundary; libtvh itself only reads those members.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

#include "types/b_field_element.hpp"
#include "types/x_field_element.hpp"

#if __has_include("table/master_table.hpp") && !defined(TVH_STANDALONE_TYPES)
#include "table/master_table.hpp"
#else
namespace triton_vm {

// (length, offset, generator): the points offset * generator^i, i < length
// (master_table.cpp:309-359)
struct ArithmeticDomain {
    size_t length = 0;
    BFieldElement offset = BFieldElement::one();
    BFieldElement generator = BFieldElement::one();

    static ArithmeticDomain of_length(size_t length) {
        if (length == 0 || (length & (length - 1)))
            throw std::invalid_argument("Domain length must be a power of 2");
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
    // half the length: generator and offset squared
    ArithmeticDomain halve() const {
        if (length % 2 != 0) throw std::invalid_argument("Cannot halve domain with odd length");
        ArithmeticDomain d = *this;
        d.length = length / 2;
        d.generator = generator * generator;
        d.offset = offset * offset;
        return d;
    }
    BFieldElement element(size_t index) const { return offset * generator.pow(index); }
    std::vector<BFieldElement> values() const {
        std::vector<BFieldElement> v;
        v.reserve(length);
        BFieldElement x = offset;
        for (size_t i = 0; i < length; i++, x *= generator) v.push_back(x);
        return v;
    }
};

}  // namespace triton_vm
#endif

#if __has_include("polynomial/polynomial.hpp") && !defined(TVH_STANDALONE_TYPES)
#include "polynomial/polynomial.hpp"
#else
namespace triton_vm {

// Dense polynomial, coefficients lowest degree first.
template <typename Field>
class Polynomial {
public:
    Polynomial() = default;
    explicit Polynomial(const std::vector<Field>& c) : coeffs_(c) {}
    explicit Polynomial(std::vector<Field>&& c) : coeffs_(std::move(c)) {}

    // coefficients without trailing zeros
    std::vector<Field> coefficients() const {
        size_t n = coeffs_.size();
        while (n > 0 && coeffs_[n - 1] == Field::zero()) n--;
        return std::vector<Field>(coeffs_.begin(), coeffs_.begin() + n);
    }
    const std::vector<Field>& raw_coefficients() const { return coeffs_; }
    std::vector<Field>& raw_coefficients() { return coeffs_; }
    // index of the highest non-zero coefficient (0 for the zero polynomial)
    size_t degree() const {
        size_t n = coefficients().size();
        return n ? n - 1 : 0;
    }
    size_t size() const { return coeffs_.size(); }
    void resize(size_t n, Field value = Field::zero()) { coeffs_.resize(n, value); }
    Field& operator[](size_t i) { return coeffs_[i]; }
    const Field& operator[](size_t i) const { return coeffs_[i]; }

    // Horner
    Field evaluate(const Field& x) const {
        Field acc = Field::zero();
        for (size_t i = coeffs_.size(); i-- > 0;) acc = acc * x + coeffs_[i];
        return acc;
    }
    template <typename EvalType>
    EvalType evaluate_at_extension(const EvalType& x) const {
        EvalType acc = EvalType::zero();
        for (size_t i = coeffs_.size(); i-- > 0;) acc = acc * x + EvalType(coeffs_[i]);
        return acc;
    }

    Polynomial operator+(const Polynomial& o) const {
        std::vector<Field> r(std::max(coeffs_.size(), o.coeffs_.size()), Field::zero());
        for (size_t i = 0; i < coeffs_.size(); i++) r[i] += coeffs_[i];
        for (size_t i = 0; i < o.coeffs_.size(); i++) r[i] += o.coeffs_[i];
        return Polynomial(std::move(r));
    }
    Polynomial operator-(const Polynomial& o) const {
        std::vector<Field> r(std::max(coeffs_.size(), o.coeffs_.size()), Field::zero());
        for (size_t i = 0; i < coeffs_.size(); i++) r[i] += coeffs_[i];
        for (size_t i = 0; i < o.coeffs_.size(); i++) r[i] -= o.coeffs_[i];
        return Polynomial(std::move(r));
    }
    Polynomial operator*(const Field& s) const {
        std::vector<Field> r(coeffs_);
        for (auto& c : r) c *= s;
        return Polynomial(std::move(r));
    }
    Polynomial operator*(const Polynomial& o) const {
        if (coeffs_.empty() || o.coeffs_.empty()) return Polynomial();
        std::vector<Field> r(coeffs_.size() + o.coeffs_.size() - 1, Field::zero());
        for (size_t i = 0; i < coeffs_.size(); i++)
            for (size_t j = 0; j < o.coeffs_.size(); j++) r[i + j] += coeffs_[i] * o.coeffs_[j];
        return Polynomial(std::move(r));
    }
    // multiply by X^n
    Polynomial shift_coefficients(size_t n) const {
        std::vector<Field> r(n, Field::zero());
        r.insert(r.end(), coeffs_.begin(), coeffs_.end());
        return Polynomial(std::move(r));
    }
    static Polynomial x_to_the(size_t n) {
        std::vector<Field> c(n + 1, Field::zero());
        c[n] = Field::one();
        return Polynomial(std::move(c));
    }
    static Polynomial from_constant(Field c) { return Polynomial(std::vector<Field>{c}); }

private:
    std::vector<Field> coeffs_;
};

using BPolynomial = Polynomial<BFieldElement>;
using XPolynomial = Polynomial<XFieldElement>;

}  // namespace triton_vm
#endif

// RandomizedLDE and LDETable with the reference's interfaces
// (include/lde/lde_randomized.hpp:17-108, include/lde/lde_table.hpp:14-57).
// The LDE and the whole commit run on the GPU; BPolynomial is the small host
// polynomial helper the zerofier functions return.
#pragma once

#include <utility>
#include <vector>

#include "triton_vm_amd/types.hpp"

namespace triton_vm {

// Coefficient vector, lowest degree first (include/polynomial/polynomial.hpp).
class BPolynomial {
public:
    BPolynomial() = default;
    explicit BPolynomial(std::vector<BFieldElement> c) : c_(std::move(c)) {}
    static BPolynomial x_to_the(size_t n) {
        std::vector<BFieldElement> c(n + 1);
        c[n] = BFieldElement::one();
        return BPolynomial(c);
    }
    size_t size() const { return c_.size(); }
    const BFieldElement& operator[](size_t i) const { return c_[i]; }
    const std::vector<BFieldElement>& raw_coefficients() const { return c_; }
    BFieldElement evaluate(BFieldElement x) const {
        BFieldElement acc;
        for (size_t i = c_.size(); i-- > 0;) acc = acc * x + c_[i];
        return acc;
    }

private:
    std::vector<BFieldElement> c_;
};

class RandomizedLDE {
public:
    // x^n - offset^n                                  (lde_randomized.cpp:10-35)
    static BPolynomial compute_zerofier(const ArithmeticDomain& domain);
    // (x^n - offset^n) * p(x)                          (lde_randomized.cpp:37-69)
    static BPolynomial mul_zerofier_with(const ArithmeticDomain& domain, const BPolynomial& p);
    // interpolant * offset^-i + zerofier * randomizer, evaluated on target_domain
    // (lde_randomized.cpp:71-154); requires randomizer length <= trace length
    static std::vector<BFieldElement> extend_column_with_randomizer(
        const std::vector<BFieldElement>& trace_column, const ArithmeticDomain& trace_domain,
        const ArithmeticDomain& target_domain, const std::vector<BFieldElement>& randomizer_coeffs);
    // XFE column, BFE randomizer lifted to (r, 0, 0)   (lde_randomized.cpp:179-300)
    static std::vector<XFieldElement> extend_xfield_column_with_randomizer(
        const std::vector<XFieldElement>& trace_column, const ArithmeticDomain& trace_domain,
        const ArithmeticDomain& quotient_domain, const std::vector<BFieldElement>& randomizer_coeffs);
    // XFE column with XFE randomizer                    (lde_randomized.cpp:302-395)
    static std::vector<XFieldElement> extend_xfield_column_with_xfield_randomizer(
        const std::vector<XFieldElement>& trace_column, const ArithmeticDomain& trace_domain,
        const ArithmeticDomain& quotient_domain, const std::vector<XFieldElement>& randomizer_coeffs);
    static bool verify_zerofier(const ArithmeticDomain& domain, const BPolynomial& zerofier);
};


class LDETable {
public:
    // row-major trace [rows][cols], randomizers [cols][k]  (lde_table.cpp:8-56)
    static std::vector<std::vector<BFieldElement>> extend_table_with_randomizers(
        const std::vector<std::vector<BFieldElement>>& trace_table,
        const ArithmeticDomain& trace_domain, const ArithmeticDomain& quotient_domain,
        const std::vector<std::vector<BFieldElement>>& randomizer_coeffs);
    // row digests and Merkle root of a row-major table  (lde_table.cpp:58-84)
    static std::pair<std::vector<Digest>, Digest> compute_merkle_tree(
        const std::vector<std::vector<BFieldElement>>& lde_table);
    // extend -> hash rows -> Merkle root                 (lde_table.cpp:86-101)
    static Digest compute_lde_and_merkle(const std::vector<std::vector<BFieldElement>>& trace_table,
                                         const ArithmeticDomain& trace_domain,
                                         const ArithmeticDomain& quotient_domain,
                                         const std::vector<std::vector<BFieldElement>>& randomizer_coeffs);
};

}  // namespace triton_vm

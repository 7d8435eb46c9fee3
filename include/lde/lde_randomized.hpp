// RandomizedLDE: drop-in for the prover's include/lde/lde_randomized.hpp:17-108.
// The column extensions run on the GPU (tvh_lde_batch_host /
// tvh_lde_xfe_batch_host); the zerofier helpers are small host polynomials.
#pragma once

#include <vector>

#include "triton_vm_amd/prover_types.hpp"
#include "types/b_field_element.hpp"
#include "types/x_field_element.hpp"

namespace triton_vm {

class RandomizedLDE {
public:
    // X^n - offset^n (lde_randomized.cpp:10-35); std::invalid_argument for length 0
    static BPolynomial compute_zerofier(const ArithmeticDomain& domain);
    // (X^n - offset^n) * p(X) (lde_randomized.cpp:37-69)
    static BPolynomial mul_zerofier_with(const ArithmeticDomain& domain,
                                         const BPolynomial& polynomial);
    // interpolant(X / offset) + zerofier * randomizer, evaluated on
    // quotient_domain (lde_randomized.cpp:71-154); randomizer length <= trace length
    static std::vector<BFieldElement> extend_column_with_randomizer(
        const std::vector<BFieldElement>& trace_column, const ArithmeticDomain& trace_domain,
        const ArithmeticDomain& quotient_domain, const std::vector<BFieldElement>& randomizer_coeffs);
    // XFE column, BFE randomizer lifted to (r, 0, 0) (lde_randomized.cpp:179-300)
    static std::vector<XFieldElement> extend_xfield_column_with_randomizer(
        const std::vector<XFieldElement>& trace_column, const ArithmeticDomain& trace_domain,
        const ArithmeticDomain& quotient_domain, const std::vector<BFieldElement>& randomizer_coeffs);
    // XFE column with XFE randomizer (lde_randomized.cpp:302-395)
    static std::vector<XFieldElement> extend_xfield_column_with_xfield_randomizer(
        const std::vector<XFieldElement>& trace_column, const ArithmeticDomain& trace_domain,
        const ArithmeticDomain& quotient_domain, const std::vector<XFieldElement>& randomizer_coeffs);
    // zerofier vanishes on every point of the domain
    static bool verify_zerofier(const ArithmeticDomain& domain, const BPolynomial& zerofier);
};

}  // namespace triton_vm

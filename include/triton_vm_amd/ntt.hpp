// NTT and LDE classes with the reference's interface (include/ntt/ntt.hpp:16-112),
// executed on the GPU through the C ABI (include/tvh.h).
// Errors: std::invalid_argument for bad sizes (as src/ntt/ntt.cpp:61-63),
// std::runtime_error for device failures.
#pragma once

#include <vector>

#include "triton_vm_amd/types.hpp"

namespace triton_vm {

class NTT {
public:
    // natural order in/out over <w_n>, w_n = 7^((p-1)/n)   (ntt.cpp:59-71)
    static void forward(std::vector<BFieldElement>& coeffs);
    // inverse including n^-1                                  (ntt.cpp:73-86)
    static void inverse(std::vector<BFieldElement>& evals);
    static std::vector<BFieldElement> interpolate(const std::vector<BFieldElement>& column);
    // zero-pad to domain_length, c_i *= offset^i, forward NTT (ntt.cpp:94-117)
    static std::vector<BFieldElement> evaluate_on_coset(const std::vector<BFieldElement>& coeffs,
                                                        size_t domain_length, BFieldElement offset);
};

class LDE {
public:
    // interpolate on the trace domain (offset ignored, as ntt.cpp:121-131) and
    // evaluate on quotient_domain
    static std::vector<BFieldElement> extend_column(const std::vector<BFieldElement>& trace_column,
                                                    const ArithmeticDomain& trace_domain,
                                                    const ArithmeticDomain& quotient_domain);
    // row-major table in, row-major table out (ntt.cpp:133-172)
    static std::vector<std::vector<BFieldElement>> extend_table(
        const std::vector<std::vector<BFieldElement>>& trace_table,
        const ArithmeticDomain& trace_domain, const ArithmeticDomain& quotient_domain);
};

namespace amd {
// throws the reference's exception type for a tvh status code
[[noreturn]] void throw_status(int status, const char* what);
inline void check(int status, const char* what) {
    if (status) throw_status(status, what);
}
// tvh_init(0) once per process (thread-safe)
void ensure_device();
}  // namespace amd

}  // namespace triton_vm

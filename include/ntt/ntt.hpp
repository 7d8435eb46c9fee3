// NTT and LDE: drop-in for the prover's include/ntt/ntt.hpp:16-112, executed
// on the GPU through the C ABI (include/tvh.h).  Same classes, static
// methods, argument meaning and exceptions: std::invalid_argument for sizes
// that are not powers of two (src/ntt/ntt.cpp:61-63), std::runtime_error for
// device failures.  Each call is synchronous and uses its own device buffers,
// so calls from several host threads at once are safe.
#pragma once

#include <cstddef>
#include <vector>

#include "triton_vm_amd/prover_types.hpp"
#include "types/b_field_element.hpp"

namespace triton_vm {

class NTT {
public:
    // out[j] = sum_i in[i] w_n^(ij), w_n = 7^((p-1)/n), natural order (ntt.cpp:59-71)
    static void forward(std::vector<BFieldElement>& coeffs);
    // the inverse transform including the factor n^-1 (ntt.cpp:73-86)
    static void inverse(std::vector<BFieldElement>& evals);
    // coefficients of the polynomial through the column's values on <w_n> (ntt.cpp:88-92)
    static std::vector<BFieldElement> interpolate(const std::vector<BFieldElement>& column);
    // values of the polynomial on offset * <w_m>, m = domain_length (ntt.cpp:94-117)
    static std::vector<BFieldElement> evaluate_on_coset(const std::vector<BFieldElement>& coeffs,
                                                        size_t domain_length,
                                                        BFieldElement offset);
};

class LDE {
public:
    // interpolate on the trace domain (its offset is not applied, as
    // ntt.cpp:121-131) and evaluate on quotient_domain
    static std::vector<BFieldElement> extend_column(const std::vector<BFieldElement>& trace_column,
                                                    const ArithmeticDomain& trace_domain,
                                                    const ArithmeticDomain& quotient_domain);
    // row-major table in, row-major table out (ntt.cpp:133-172)
    static std::vector<std::vector<BFieldElement>> extend_table(
        const std::vector<std::vector<BFieldElement>>& trace_table,
        const ArithmeticDomain& trace_domain, const ArithmeticDomain& quotient_domain);
};

namespace amd {
// throws the prover's exception type for a tvh status code
[[noreturn]] void throw_status(int status, const char* what);
inline void check(int status, const char* what) {
    if (status) throw_status(status, what);
}
// tvh_init(0) once per process (thread-safe)
void ensure_device();
}  // namespace amd

}  // namespace triton_vm

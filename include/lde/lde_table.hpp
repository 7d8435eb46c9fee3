// LDETable: drop-in for the prover's include/lde/lde_table.hpp:14-57.  The
// whole table goes to the GPU in one call (extension, row hashing and the
// Merkle tree in tvh_commit_host / tvh_lde_batch_host).
#pragma once

#include <utility>
#include <vector>

#include "hash/tip5.hpp"
#include "triton_vm_amd/prover_types.hpp"
#include "types/b_field_element.hpp"
#include "types/digest.hpp"

namespace triton_vm {

class LDETable {
public:
    // row-major trace [rows][cols], randomizers [cols][k] (lde_table.cpp:8-56)
    static std::vector<std::vector<BFieldElement>> extend_table_with_randomizers(
        const std::vector<std::vector<BFieldElement>>& trace_table,
        const ArithmeticDomain& trace_domain, const ArithmeticDomain& quotient_domain,
        const std::vector<std::vector<BFieldElement>>& randomizer_coeffs);
    // row digests and Merkle root of a row-major table (lde_table.cpp:58-84)
    static std::pair<std::vector<Digest>, Digest> compute_merkle_tree(
        const std::vector<std::vector<BFieldElement>>& lde_table);
    // extend -> hash rows -> Merkle root (lde_table.cpp:86-101)
    static Digest compute_lde_and_merkle(
        const std::vector<std::vector<BFieldElement>>& trace_table,
        const ArithmeticDomain& trace_domain, const ArithmeticDomain& quotient_domain,
        const std::vector<std::vector<BFieldElement>>& randomizer_coeffs);
};

}  // namespace triton_vm

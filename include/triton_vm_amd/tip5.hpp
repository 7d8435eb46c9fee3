// Tip5 with the reference's interface (include/hash/tip5.hpp:18-68).  Every
// permutation runs on the GPU (one state per call here; batches go through
// hash_rows / hash_pairs / permutation_batch, which amortise the launch).
#pragma once

#include <array>
#include <vector>

#include "triton_vm_amd/types.hpp"

namespace triton_vm {

class Tip5 {
public:
    static constexpr size_t STATE_SIZE = 16;
    static constexpr size_t NUM_SPLIT_AND_LOOKUP = 4;
    static constexpr size_t CAPACITY = 6;
    static constexpr size_t RATE = 10;
    static constexpr size_t NUM_ROUNDS = 5;

    std::array<BFieldElement, STATE_SIZE> state;

    Tip5() { state.fill(BFieldElement::zero()); }
    static Tip5 init() { return Tip5(); }  // zero state (tip5.cpp:85-90)

    void permutation();
    // state before and after each round (tip5.cpp:98-113)
    std::vector<std::array<BFieldElement, STATE_SIZE>> trace();
    // overwrite absorb of up to RATE elements, then permute (tip5.cpp:205-212)
    void absorb(const std::vector<BFieldElement>& elements);
    std::vector<BFieldElement> squeeze(size_t count);

    static Digest hash_10(const std::array<BFieldElement, RATE>& input);
    static Digest hash_varlen(const std::vector<BFieldElement>& input);
    static Digest hash_pair(const Digest& left, const Digest& right);

    // batch forms (no reference counterpart): one launch for many inputs
    static void permutation_batch(std::vector<std::array<BFieldElement, STATE_SIZE>>& states);
    static std::vector<Digest> hash_pairs(const std::vector<Digest>& left,
                                          const std::vector<Digest>& right);
    // hash_varlen of every row of a row-major table
    static std::vector<Digest> hash_rows(const std::vector<std::vector<BFieldElement>>& rows);
};

}  // namespace triton_vm

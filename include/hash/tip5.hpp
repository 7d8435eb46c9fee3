// Tip5: drop-in for the prover's include/hash/tip5.hpp:18-68.  Every
// permutation runs on the GPU (one state per call for the instance methods;
// the batch forms below amortise the launch over many inputs).  The public
// constant tables are the specification's (src/hash/tip5.cpp:7-79).
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "types/b_field_element.hpp"
#include "types/digest.hpp"

namespace triton_vm {

class Tip5 {
public:
    static constexpr size_t STATE_SIZE = 16;
    static constexpr size_t NUM_SPLIT_AND_LOOKUP = 4;
    static constexpr size_t CAPACITY = 6;
    static constexpr size_t RATE = 10;
    static constexpr size_t NUM_ROUNDS = 5;

    // S-box of the split-and-lookup lanes: ((x + 1)^3 mod 257) - 1
    static const std::array<uint8_t, 256> LOOKUP_TABLE;
    // first column of the circulant MDS matrix (SHA-256("Tip5") as 16 LE u16)
    static const std::array<int64_t, STATE_SIZE> MDS_MATRIX_FIRST_COLUMN;
    static const std::array<BFieldElement, NUM_ROUNDS * STATE_SIZE> ROUND_CONSTANTS;

    std::array<BFieldElement, STATE_SIZE> state;

    Tip5();
    // the variable-length sponge's zero state (tip5.cpp:85-90)
    static Tip5 init();

    void permutation();
    // round round_index (< NUM_ROUNDS) of the permutation: S-box layer, MDS,
    // round constants (tip5.cpp:115-125); std::out_of_range beyond the last round
    void round(size_t round_index);

    // overwrite the rate with up to RATE elements, then permute (tip5.cpp:205-212)
    void absorb(const std::vector<BFieldElement>& elements);
    // the rate's first count (<= RATE) elements, then permute
    std::vector<BFieldElement> squeeze(size_t count);

    static Digest hash_10(const std::array<BFieldElement, RATE>& input);
    static Digest hash_varlen(const std::vector<BFieldElement>& input);
    static Digest hash_pair(const Digest& left, const Digest& right);

    // the state before the first round and after each round (tip5.cpp:98-113);
    // leaves the state permuted, as the prover's does
    std::vector<std::array<BFieldElement, STATE_SIZE>> trace();

    // Batch forms (no counterpart in the prover's class): one launch for many
    // inputs.
    static void permutation_batch(std::vector<std::array<BFieldElement, STATE_SIZE>>& states);
    static std::vector<Digest> hash_pairs(const std::vector<Digest>& left,
                                          const std::vector<Digest>& right);
    // hash_varlen of every row of a row-major table
    static std::vector<Digest> hash_rows(const std::vector<std::vector<BFieldElement>>& rows);
};

}  // namespace triton_vm

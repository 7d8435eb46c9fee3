// MerkleTree with the reference's interface (include/merkle/merkle_tree.hpp:16-99).
// The tree is built on the GPU (heap layout, node i = hash_pair(2i, 2i+1)) and
// kept on the host for paths; verification hashes on the GPU.
#pragma once

#include <vector>

#include "triton_vm_amd/types.hpp"

namespace triton_vm {

class MerkleTree {
public:
    // throws std::invalid_argument for no leaves / non-power-of-two (merkle_tree.cpp:12-33)
    explicit MerkleTree(const std::vector<Digest>& leaves);
    static MerkleTree parallel_new(const std::vector<Digest>& leaves) { return MerkleTree(leaves); }

    Digest root() const { return nodes_[1]; }
    size_t num_leaves() const { return num_leaves_; }
    size_t height() const;
    Digest leaf(size_t index) const;                                  // std::out_of_range
    std::vector<Digest> authentication_path(size_t leaf_index) const;  // bottom-up siblings
    // needed minus computable nodes, sorted descending (merkle_tree.cpp:115-226)
    std::vector<Digest> authentication_structure(const std::vector<size_t>& leaf_indices) const;

    static bool verify(const Digest& root, size_t leaf_index, const Digest& leaf,
                       const std::vector<Digest>& auth_path);
    static bool verify_authentication_structure(const Digest& root, size_t num_leaves,
                                                const std::vector<size_t>& leaf_indices,
                                                const std::vector<Digest>& leaves,
                                                const std::vector<Digest>& auth_structure);
    static size_t sibling_index(size_t node_index) { return node_index ^ 1; }
    static size_t parent_index(size_t node_index) { return node_index / 2; }

    // heap nodes, index 1 = root (no reference counterpart: exposes the buffer)
    const std::vector<Digest>& nodes() const { return nodes_; }

private:
    std::vector<Digest> nodes_;
    size_t num_leaves_ = 0;
};

// hash_varlen of a row; the empty row hashes to the zero digest (merkle_tree.cpp:366-373)
Digest hash_row(const std::vector<BFieldElement>& row);
// table_commitment.cpp:61-77
Digest hash_bfield_row(const std::vector<BFieldElement>& row);
Digest hash_xfield_row(const std::vector<XFieldElement>& row);

}  // namespace triton_vm

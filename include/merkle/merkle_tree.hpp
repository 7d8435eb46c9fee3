// MerkleTree: drop-in for the prover's include/merkle/merkle_tree.hpp:16-99.
// The tree is built on the GPU in the prover's heap layout (node i =
// hash_pair(node 2i, node 2i+1), root at 1, leaves at [n, 2n)) and kept on
// the host for paths and structures; verification hashes on the GPU.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <vector>

#include "hash/tip5.hpp"
#include "types/b_field_element.hpp"
#include "types/digest.hpp"

namespace triton_vm {

class MerkleTree {
public:
    // std::invalid_argument for no leaves or a count that is not a power of
    // two (merkle_tree.cpp:12-33)
    explicit MerkleTree(const std::vector<Digest>& leaves);
    static MerkleTree parallel_new(const std::vector<Digest>& leaves);

    Digest root() const;
    size_t num_leaves() const { return num_leaves_; }
    size_t height() const;
    Digest leaf(size_t index) const;  // std::out_of_range
    // bottom-up siblings of the leaf's path (merkle_tree.cpp:93-113)
    std::vector<Digest> authentication_path(size_t leaf_index) const;
    // nodes needed to verify all the leaves minus the nodes they let a
    // verifier compute, by heap index descending (merkle_tree.cpp:115-265)
    std::vector<Digest> authentication_structure(const std::vector<size_t>& leaf_indices) const;

    static bool verify(const Digest& root, size_t leaf_index, const Digest& leaf,
                       const std::vector<Digest>& auth_path);
    static bool verify_authentication_structure(const Digest& root, size_t num_leaves,
                                                const std::vector<size_t>& leaf_indices,
                                                const std::vector<Digest>& leaves,
                                                const std::vector<Digest>& auth_structure);

    static size_t sibling_index(size_t node_index) { return node_index ^ 1; }
    static size_t parent_index(size_t node_index) { return node_index / 2; }

    // heap nodes, index 1 = root, index 0 unused (no counterpart in the
    // prover's class, whose node storage is private)
    const std::vector<Digest>& nodes() const { return nodes_; }
    // a tree from heap nodes built elsewhere (e.g. on the device):
    // nodes.size() == 2 * num_leaves
    static MerkleTree from_heap_nodes(std::vector<Digest> nodes);

private:
    MerkleTree() = default;
    std::vector<Digest> nodes_;
    size_t num_leaves_ = 0;
};

// Tip5::hash_varlen of the row; the empty row hashes to the zero digest
// (merkle_tree.cpp:366-373)
Digest hash_row(const std::vector<BFieldElement>& row);

}  // namespace triton_vm

// TableCommitment: drop-in for the prover's include/table/table_commitment.hpp:19-72
// (src/table/table_commitment.cpp:9-77).  The row digests and the Merkle tree
// are computed on the GPU; the tree's nodes are kept on the host for paths.
//
// commit(MasterMainTable / MasterAuxTable) needs the prover's master table
// classes, so it is defined (inline, below) only where the prover's
// table/master_table.hpp is on the include path; from_digests and the
// row-table forms are always available.
#pragma once

#include <cstddef>
#include <memory>
#include <vector>

#include "merkle/merkle_tree.hpp"
#include "triton_vm_amd/prover_types.hpp"
#include "types/b_field_element.hpp"
#include "types/digest.hpp"
#include "types/x_field_element.hpp"

#if __has_include("table/master_table.hpp") && !defined(TVH_STANDALONE_TYPES)
#define TVH_HAVE_MASTER_TABLES 1
#else
#define TVH_HAVE_MASTER_TABLES 0
namespace triton_vm {
class MasterMainTable;
class MasterAuxTable;
}  // namespace triton_vm
#endif

namespace triton_vm {

class TableCommitment {
public:
    // hash_bfield_row of every row, then the Merkle tree over the digests
    static TableCommitment commit(const MasterMainTable& table);
    // hash_xfield_row of every row, then the Merkle tree
    static TableCommitment commit(const MasterAuxTable& table);
    // the Merkle tree over precomputed row digests (a power-of-two count)
    static TableCommitment from_digests(const std::vector<Digest>& row_digests);

    // Row-table forms behind commit(), one GPU call each (no counterpart in
    // the prover's class): rows[i] is row i; every row has the same width.
    static TableCommitment commit_bfield_rows(const std::vector<std::vector<BFieldElement>>& rows);
    static TableCommitment commit_xfield_rows(const std::vector<std::vector<XFieldElement>>& rows);
    // the same over row pointers (rows[i] points at `width` elements)
    static TableCommitment commit_bfield_rows(const BFieldElement* const* rows, size_t num_rows,
                                              size_t width);
    static TableCommitment commit_xfield_rows(const XFieldElement* const* rows, size_t num_rows,
                                              size_t width);

    Digest root() const;
    size_t num_rows() const;
    std::vector<Digest> authentication_path(size_t row_index) const;
    std::vector<Digest> authentication_structure(const std::vector<size_t>& row_indices) const;
    static bool verify(const Digest& root, size_t row_index, const Digest& row_digest,
                       const std::vector<Digest>& auth_path);
    const MerkleTree& tree() const { return *tree_; }

private:
    std::unique_ptr<MerkleTree> tree_;
    explicit TableCommitment(std::unique_ptr<MerkleTree> tree);
};

// Tip5::hash_varlen of the row (table_commitment.cpp:61-63)
Digest hash_bfield_row(const std::vector<BFieldElement>& row);
// Tip5::hash_varlen of the row flattened to (c0, c1, c2) per element
// (table_commitment.cpp:65-77)
Digest hash_xfield_row(const std::vector<XFieldElement>& row);

#if TVH_HAVE_MASTER_TABLES
inline TableCommitment TableCommitment::commit(const MasterMainTable& table) {
    std::vector<const BFieldElement*> rows(table.num_rows());
    size_t width = table.num_rows() ? table.row(0).size() : 0;
    for (size_t i = 0; i < rows.size(); i++) {
        const auto& r = table.row(i);
        if (r.size() != width) throw std::invalid_argument("TableCommitment::commit: ragged table");
        rows[i] = r.data();
    }
    return commit_bfield_rows(rows.data(), rows.size(), width);
}
inline TableCommitment TableCommitment::commit(const MasterAuxTable& table) {
    std::vector<const XFieldElement*> rows(table.num_rows());
    size_t width = table.num_rows() ? table.row(0).size() : 0;
    for (size_t i = 0; i < rows.size(); i++) {
        const auto& r = table.row(i);
        if (r.size() != width) throw std::invalid_argument("TableCommitment::commit: ragged table");
        rows[i] = r.data();
    }
    return commit_xfield_rows(rows.data(), rows.size(), width);
}
#endif

}  // namespace triton_vm

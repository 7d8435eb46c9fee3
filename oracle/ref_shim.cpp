// C-ABI shim over the reference prover's own CPU classes, compiled together
// with the reference sources under /root/reference by oracle/Makefile into
// oracle/_ref/libtvm_ref.so.  TEST INFRASTRUCTURE ONLY: used to validate the
// C oracle, to write golden fixtures (tests/golden/make_golden.py) and as the
// "reference" CPU baseline in bench.py.  The shipped library never loads it.
//
// Entry points wrap, without changing behaviour:
//   NTT::forward / NTT::inverse                      src/ntt/ntt.cpp:59-86
//   RandomizedLDE::extend_column_with_randomizer     src/lde/lde_randomized.cpp:71-154
//   RandomizedLDE::extend_xfield_column_with_randomizer   :179-300
//   Tip5::permutation / hash_varlen / hash_pair      src/hash/tip5.cpp:92-297
//   MerkleTree (+ authentication_structure)          src/merkle/merkle_tree.cpp:12-226
//   LDETable::compute_lde_and_merkle pieces          src/lde/lde_table.cpp:8-101
#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

#include <omp.h>

#include "hash/tip5.hpp"
#include "lde/lde_randomized.hpp"
#include "lde/lde_table.hpp"
#include "merkle/merkle_tree.hpp"
#include "ntt/ntt.hpp"

using namespace triton_vm;

namespace {
std::vector<BFieldElement> to_bfe(const uint64_t* p, size_t n) {
    std::vector<BFieldElement> v(n);
    for (size_t i = 0; i < n; i++) v[i] = BFieldElement(p[i]);
    return v;
}
Digest to_digest(const uint64_t* p) {
    return Digest(BFieldElement(p[0]), BFieldElement(p[1]), BFieldElement(p[2]),
                  BFieldElement(p[3]), BFieldElement(p[4]));
}
void from_digest(const Digest& d, uint64_t* p) {
    for (size_t i = 0; i < 5; i++) p[i] = d[i].value();
}
ArithmeticDomain domain(size_t len, uint64_t off) {
    ArithmeticDomain d;
    d.length = len;
    d.offset = BFieldElement(off);
    uint32_t lg = 0;
    while ((size_t(1) << lg) < len) lg++;
    d.generator = BFieldElement::primitive_root_of_unity(lg);
    return d;
}
}  // namespace

extern "C" {

void ref_set_threads(int t) { omp_set_num_threads(t); }

int ref_ntt_forward(uint64_t* a, size_t n) {
    try {
        auto v = to_bfe(a, n);
        NTT::forward(v);
        for (size_t i = 0; i < n; i++) a[i] = v[i].value();
        return 0;
    } catch (const std::exception&) { return -1; }
}

int ref_ntt_inverse(uint64_t* a, size_t n) {
    try {
        auto v = to_bfe(a, n);
        NTT::inverse(v);
        for (size_t i = 0; i < n; i++) a[i] = v[i].value();
        return 0;
    } catch (const std::exception&) { return -1; }
}

int ref_lde_column(const uint64_t* trace, size_t n, uint64_t toff, const uint64_t* rand,
                   size_t k, uint64_t goff, size_t m, uint64_t* out) {
    try {
        auto col = to_bfe(trace, n);
        std::vector<BFieldElement> r = rand ? to_bfe(rand, k) : std::vector<BFieldElement>{};
        auto res = RandomizedLDE::extend_column_with_randomizer(col, domain(n, toff),
                                                                domain(m, goff), r);
        if (res.size() != m) return -2;
        for (size_t i = 0; i < m; i++) out[i] = res[i].value();
        return 0;
    } catch (const std::exception&) { return -1; }
}

// XFE column in limb-interleaved form (3*n values, x[3*i + c]); BFE randomizer
// of length k (lifted to limb 0), or an XFE randomizer of 3*k values when
// xfe_rand != 0.  Output 3*m values, interleaved the same way.
int ref_lde_xfe_column(const uint64_t* trace3, size_t n, uint64_t toff, const uint64_t* rand,
                       size_t k, int xfe_rand, uint64_t goff, size_t m, uint64_t* out3) {
    try {
        std::vector<XFieldElement> col(n);
        for (size_t i = 0; i < n; i++)
            col[i] = XFieldElement(BFieldElement(trace3[3 * i]), BFieldElement(trace3[3 * i + 1]),
                                   BFieldElement(trace3[3 * i + 2]));
        std::vector<XFieldElement> res;
        if (xfe_rand) {
            std::vector<XFieldElement> r(k);
            for (size_t i = 0; i < k; i++)
                r[i] = XFieldElement(BFieldElement(rand[3 * i]), BFieldElement(rand[3 * i + 1]),
                                     BFieldElement(rand[3 * i + 2]));
            res = RandomizedLDE::extend_xfield_column_with_xfield_randomizer(
                col, domain(n, toff), domain(m, goff), r);
        } else {
            std::vector<BFieldElement> r = rand ? to_bfe(rand, k) : std::vector<BFieldElement>{};
            res = RandomizedLDE::extend_xfield_column_with_randomizer(col, domain(n, toff),
                                                                      domain(m, goff), r);
        }
        if (res.size() != m) return -2;
        for (size_t i = 0; i < m; i++)
            for (size_t c = 0; c < 3; c++) out3[3 * i + c] = res[i].coeff(c).value();
        return 0;
    } catch (const std::exception&) { return -1; }
}

void ref_tip5_permutation(uint64_t* st) {
    Tip5 t;
    for (size_t i = 0; i < 16; i++) t.state[i] = BFieldElement(st[i]);
    t.permutation();
    for (size_t i = 0; i < 16; i++) st[i] = t.state[i].value();
}

// state before and after each round (Tip5::trace): out holds 6 x 16 u64
void ref_tip5_trace(const uint64_t* st, uint64_t* out) {
    Tip5 t;
    for (size_t i = 0; i < 16; i++) t.state[i] = BFieldElement(st[i]);
    auto tr = t.trace();
    for (size_t j = 0; j < tr.size() && j < 6; j++)
        for (size_t i = 0; i < 16; i++) out[16 * j + i] = tr[j][i].value();
}

void ref_hash_varlen(const uint64_t* in, size_t len, uint64_t* out) {
    from_digest(Tip5::hash_varlen(to_bfe(in, len)), out);
}

void ref_hash_pair(const uint64_t* l, const uint64_t* r, uint64_t* out) {
    from_digest(Tip5::hash_pair(to_digest(l), to_digest(r)), out);
}

// Row digests of a column-major table (LDETable::compute_merkle_tree's loop).
void ref_hash_rows_colmajor(const uint64_t* t, size_t rows, size_t cols, uint64_t* dg) {
    std::vector<BFieldElement> row(cols);
    for (size_t r = 0; r < rows; r++) {
        for (size_t c = 0; c < cols; c++) row[c] = BFieldElement(t[c * rows + r]);
        from_digest(Tip5::hash_varlen(row), dg + 5 * r);
    }
}

// Root (and optionally the heap nodes 1..2n-1 as 5*2n values, node 0 zero).
int ref_merkle(const uint64_t* leaves, size_t n, uint64_t* root, uint64_t* heap_nodes) {
    try {
        std::vector<Digest> lv(n);
        for (size_t i = 0; i < n; i++) lv[i] = to_digest(leaves + 5 * i);
        MerkleTree t(lv);
        from_digest(t.root(), root);
        if (heap_nodes) {
            std::memset(heap_nodes, 0, 5 * sizeof(uint64_t));
            for (size_t i = 0; i < n; i++) from_digest(t.leaf(i), heap_nodes + 5 * (n + i));
            // internal nodes through authentication paths of leaf pairs
            for (size_t i = 0; i < n; i++) {
                auto path = t.authentication_path(i);
                size_t v = n + i;
                for (const auto& d : path) {
                    from_digest(d, heap_nodes + 5 * (v ^ 1));
                    v >>= 1;
                }
            }
            from_digest(t.root(), heap_nodes + 5);
        }
        return 0;
    } catch (const std::exception&) { return -1; }
}

// authentication_structure digests (5 per node); returns the node count.
long ref_auth_structure(const uint64_t* leaves, size_t n, const size_t* idx, size_t cnt,
                        uint64_t* out, size_t cap) {
    try {
        std::vector<Digest> lv(n);
        for (size_t i = 0; i < n; i++) lv[i] = to_digest(leaves + 5 * i);
        MerkleTree t(lv);
        std::vector<size_t> ix(idx, idx + cnt);
        auto s = t.authentication_structure(ix);
        for (size_t i = 0; i < s.size() && i < cap; i++) from_digest(s[i], out + 5 * i);
        return (long)s.size();
    } catch (const std::exception&) { return -1; }
}

// Full commit of a column-major trace: LDE per column, row hashing, Merkle root.
int ref_commit(const uint64_t* trace, size_t n, size_t cols, uint64_t toff, const uint64_t* rand,
               size_t k, uint64_t goff, size_t m, uint64_t* codewords, uint64_t* digests,
               uint64_t* root) {
    try {
        std::vector<std::vector<BFieldElement>> table(n, std::vector<BFieldElement>(cols));
        for (size_t c = 0; c < cols; c++)
            for (size_t r = 0; r < n; r++) table[r][c] = BFieldElement(trace[c * n + r]);
        std::vector<std::vector<BFieldElement>> rnd(cols);
        for (size_t c = 0; c < cols; c++) rnd[c] = rand ? to_bfe(rand + c * k, k)
                                                         : std::vector<BFieldElement>{};
        auto lde = LDETable::extend_table_with_randomizers(table, domain(n, toff),
                                                           domain(m, goff), rnd);
        if (codewords)
            for (size_t r = 0; r < m; r++)
                for (size_t c = 0; c < cols; c++) codewords[c * m + r] = lde[r][c].value();
        auto [hashes, rt] = LDETable::compute_merkle_tree(lde);
        if (digests)
            for (size_t r = 0; r < m; r++) from_digest(hashes[r], digests + 5 * r);
        from_digest(rt, root);
        return 0;
    } catch (const std::exception&) { return -1; }
}

// The same commit with the work spread over the host's cores: the reference's
// per-column LDE (RandomizedLDE::extend_column_with_randomizer, lde_randomized.cpp:71-154)
// in an OpenMP loop over columns, Tip5::hash_varlen over rows in an OpenMP
// loop, then MerkleTree (whose levels are OpenMP loops, merkle_tree.cpp:41-61):
// the parallel form main_gpu.cpp:113-123 sets its OpenMP threads up for.
int ref_commit_parallel(const uint64_t* trace, size_t n, size_t cols, uint64_t toff, size_t m,
                        uint64_t goff, uint64_t* root) {
    try {
        std::vector<std::vector<BFieldElement>> lde(cols);
        const ArithmeticDomain td = domain(n, toff), qd = domain(m, goff);
        int bad = 0;
#pragma omp parallel for schedule(dynamic, 1)
        for (long c = 0; c < (long)cols; c++) {
            try {
                lde[c] = RandomizedLDE::extend_column_with_randomizer(to_bfe(trace + c * n, n), td, qd,
                                                                       {});
            } catch (const std::exception&) {
#pragma omp atomic write
                bad = 1;
            }
        }
        if (bad) return -1;
        std::vector<Digest> leaves(m);
#pragma omp parallel for schedule(static)
        for (long r = 0; r < (long)m; r++) {
            std::vector<BFieldElement> row(cols);
            for (size_t c = 0; c < cols; c++) row[c] = lde[c][r];
            leaves[r] = Tip5::hash_varlen(row);
        }
        MerkleTree t(leaves);
        from_digest(t.root(), root);
        return 0;
    } catch (const std::exception&) { return -1; }
}

}  // extern "C"

// Drop-in implementations of the reference's NTT / LDE / RandomizedLDE /
// LDETable / Tip5 / MerkleTree / TableCommitment classes over the C ABI of
// libtvh.  Each method forwards its bulk work to one tvh_*_host call; the host
// side only converts between the reference's containers and flat column-major
// u64 buffers.
#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>

#include "hash/tip5.hpp"
#include "lde/lde_randomized.hpp"
#include "lde/lde_table.hpp"
#include "merkle/merkle_tree.hpp"
#include "ntt/ntt.hpp"
#include "table/table_commitment.hpp"
#include "tip5_constants.hpp"
#include "tvh.h"

namespace triton_vm {

namespace amd {

void throw_status(int status, const char* what) {
    std::string msg = std::string(what) + ": " + tvh_status_string(status) + " (" +
                      tvh_last_error() + ")";
    switch (status) {
        case TVH_ERR_INVALID_ARG:
        case TVH_ERR_NOT_POW2:
        case TVH_ERR_NOT_CANONICAL:
        case TVH_ERR_UNSUPPORTED:
        case TVH_ERR_LAYOUT:
            throw std::invalid_argument(msg);
        case TVH_ERR_RANGE:
            throw std::out_of_range(msg);
        default:
            throw std::runtime_error(msg);
    }
}

void ensure_device() {
    static std::once_flag once;
    static int rc = 0;
    std::call_once(once, [] { rc = tvh_init(0); });
    if (rc) throw_status(rc, "tvh_init");
}

}  // namespace amd

namespace {

static_assert(sizeof(BFieldElement) == sizeof(uint64_t), "BFieldElement must be one u64");
static_assert(sizeof(Digest) == 5 * sizeof(uint64_t), "Digest must be 5 u64");

uint64_t* raw(std::vector<BFieldElement>& v) { return reinterpret_cast<uint64_t*>(v.data()); }
const uint64_t* raw(const std::vector<BFieldElement>& v) {
    return reinterpret_cast<const uint64_t*>(v.data());
}

bool pow2(size_t n) { return n && !(n & (n - 1)); }

std::vector<uint64_t> to_colmajor(const std::vector<std::vector<BFieldElement>>& rows) {
    size_t n = rows.size(), w = n ? rows[0].size() : 0;
    std::vector<uint64_t> out(n * w);
    for (size_t r = 0; r < n; r++) {
        if (rows[r].size() != w) throw std::invalid_argument("ragged table");
        for (size_t c = 0; c < w; c++) out[c * n + r] = rows[r][c].value();
    }
    return out;
}

std::vector<std::vector<BFieldElement>> to_rowmajor(const std::vector<uint64_t>& cm, size_t n,
                                                    size_t w) {
    std::vector<std::vector<BFieldElement>> rows(n, std::vector<BFieldElement>(w));
    for (size_t c = 0; c < w; c++)
        for (size_t r = 0; r < n; r++) rows[r][c] = BFieldElement(cm[c * n + r]);
    return rows;
}

std::vector<uint64_t> bfe_rand_rows(const std::vector<std::vector<BFieldElement>>& rnd, size_t cols,
                                    size_t* k_out) {
    size_t k = cols && rnd.size() == cols ? rnd[0].size() : 0;
    for (const auto& r : rnd)
        if (r.size() != k) throw std::invalid_argument("randomizer lengths differ across columns");
    std::vector<uint64_t> out(cols * k);
    for (size_t c = 0; c < cols && k; c++)
        for (size_t i = 0; i < k; i++) out[c * k + i] = rnd[c][i].value();
    *k_out = k;
    return out;
}

}  // namespace

// ------------------------------------------------------------------ NTT / LDE
void NTT::forward(std::vector<BFieldElement>& coeffs) {
    if (!pow2(coeffs.size())) throw std::invalid_argument("NTT size must be a power of 2");
    amd::ensure_device();
    amd::check(tvh_ntt_forward_host(raw(coeffs), coeffs.size()), "NTT::forward");
}

void NTT::inverse(std::vector<BFieldElement>& evals) {
    if (!pow2(evals.size())) throw std::invalid_argument("NTT size must be a power of 2");
    amd::ensure_device();
    amd::check(tvh_ntt_inverse_host(raw(evals), evals.size()), "NTT::inverse");
}

std::vector<BFieldElement> NTT::interpolate(const std::vector<BFieldElement>& column) {
    std::vector<BFieldElement> c = column;
    inverse(c);
    return c;
}

std::vector<BFieldElement> NTT::evaluate_on_coset(const std::vector<BFieldElement>& coeffs,
                                                  size_t m, BFieldElement offset) {
    std::vector<BFieldElement> ext(m);
    BFieldElement s = BFieldElement::one();
    for (size_t i = 0; i < m; i++) {
        ext[i] = (i < coeffs.size() ? coeffs[i] : BFieldElement::zero()) * s;
        s *= offset;
    }
    forward(ext);
    return ext;
}

std::vector<BFieldElement> LDE::extend_column(const std::vector<BFieldElement>& trace_column,
                                              const ArithmeticDomain& /*trace_domain*/,
                                              const ArithmeticDomain& quotient_domain) {
    size_t n = trace_column.size(), m = quotient_domain.length;
    amd::ensure_device();
    std::vector<BFieldElement> out(m);
    // the reference interpolates without the trace offset (ntt.cpp:121-131)
    amd::check(tvh_lde_batch_host(raw(trace_column), 1, n, nullptr, 0, 1,
                                  quotient_domain.offset.value(), m, raw(out)),
               "LDE::extend_column");
    return out;
}

std::vector<std::vector<BFieldElement>> LDE::extend_table(
    const std::vector<std::vector<BFieldElement>>& table, const ArithmeticDomain& trace_domain,
    const ArithmeticDomain& quotient_domain) {
    if (table.empty()) return {};
    // (only the domains' members are read here: the prover's own
    // ArithmeticDomain methods are defined in its master_table.cpp)
    ArithmeticDomain td = trace_domain;
    td.offset = BFieldElement::one();
    return LDETable::extend_table_with_randomizers(
        table, td, quotient_domain,
        std::vector<std::vector<BFieldElement>>(table[0].size()));
}

// ------------------------------------------------------------------ RandomizedLDE
BPolynomial RandomizedLDE::compute_zerofier(const ArithmeticDomain& d) {
    if (d.length == 0) throw std::invalid_argument("Domain length must be > 0");
    std::vector<BFieldElement> c(d.length + 1);
    c[d.length] = BFieldElement::one();
    c[0] = -d.offset.pow(d.length);
    return BPolynomial(c);
}

BPolynomial RandomizedLDE::mul_zerofier_with(const ArithmeticDomain& d, const BPolynomial& p) {
    BFieldElement on = d.offset.pow(d.length);
    std::vector<BFieldElement> c(std::max(p.size() + d.length, p.size()));
    for (size_t i = 0; i < p.size(); i++) {
        c[i + d.length] += p[i];
        c[i] -= p[i] * on;
    }
    return BPolynomial(c);
}

bool RandomizedLDE::verify_zerofier(const ArithmeticDomain& d, const BPolynomial& z) {
    BFieldElement x = d.offset;
    for (size_t i = 0; i < d.length; i++, x *= d.generator)
        if (!z.evaluate(x).is_zero()) return false;
    return true;
}

std::vector<BFieldElement> RandomizedLDE::extend_column_with_randomizer(
    const std::vector<BFieldElement>& col, const ArithmeticDomain& td, const ArithmeticDomain& gd,
    const std::vector<BFieldElement>& r) {
    amd::ensure_device();
    std::vector<BFieldElement> out(gd.length);
    amd::check(tvh_lde_batch_host(raw(col), 1, col.size(), r.empty() ? nullptr : raw(r), r.size(),
                                  td.offset.value(), gd.offset.value(), gd.length, raw(out)),
               "RandomizedLDE::extend_column_with_randomizer");
    return out;
}

namespace {
std::vector<XFieldElement> xfe_lde(const std::vector<XFieldElement>& col,
                                   const ArithmeticDomain& td, const ArithmeticDomain& gd,
                                   const std::vector<uint64_t>& r, size_t k, int xfe_rand) {
    amd::ensure_device();
    std::vector<uint64_t> in(3 * col.size()), out(3 * gd.length);
    for (size_t i = 0; i < col.size(); i++)
        for (size_t j = 0; j < 3; j++) in[3 * i + j] = col[i].coeff(j).value();
    amd::check(tvh_lde_xfe_batch_host(in.data(), 1, col.size(), k ? r.data() : nullptr, k,
                                      td.offset.value(), gd.offset.value(), gd.length, out.data(),
                                      xfe_rand),
               "RandomizedLDE::extend_xfield_column");
    std::vector<XFieldElement> res(gd.length);
    for (size_t i = 0; i < gd.length; i++)
        res[i] = XFieldElement(BFieldElement(out[3 * i]), BFieldElement(out[3 * i + 1]),
                               BFieldElement(out[3 * i + 2]));
    return res;
}
}  // namespace

std::vector<XFieldElement> RandomizedLDE::extend_xfield_column_with_randomizer(
    const std::vector<XFieldElement>& col, const ArithmeticDomain& td, const ArithmeticDomain& qd,
    const std::vector<BFieldElement>& r) {
    std::vector<uint64_t> rr(r.size());
    for (size_t i = 0; i < r.size(); i++) rr[i] = r[i].value();
    return xfe_lde(col, td, qd, rr, r.size(), 0);
}

std::vector<XFieldElement> RandomizedLDE::extend_xfield_column_with_xfield_randomizer(
    const std::vector<XFieldElement>& col, const ArithmeticDomain& td, const ArithmeticDomain& qd,
    const std::vector<XFieldElement>& r) {
    std::vector<uint64_t> rr(3 * r.size());
    for (size_t i = 0; i < r.size(); i++)
        for (size_t j = 0; j < 3; j++) rr[3 * i + j] = r[i].coeff(j).value();
    return xfe_lde(col, td, qd, rr, r.size(), 1);
}

// ------------------------------------------------------------------ LDETable
std::vector<std::vector<BFieldElement>> LDETable::extend_table_with_randomizers(
    const std::vector<std::vector<BFieldElement>>& table, const ArithmeticDomain& td,
    const ArithmeticDomain& qd, const std::vector<std::vector<BFieldElement>>& rnd) {
    if (table.empty()) return {};
    size_t n = table.size(), w = table[0].size(), m = qd.length;
    if (rnd.size() != w)
        throw std::invalid_argument("Randomizer coefficients count must match column count");
    size_t k = 0;
    std::vector<uint64_t> r = bfe_rand_rows(rnd, w, &k);
    std::vector<uint64_t> cm = to_colmajor(table), out(m * w);
    amd::ensure_device();
    amd::check(tvh_lde_batch_host(cm.data(), w, n, k ? r.data() : nullptr, k, td.offset.value(),
                                  qd.offset.value(), m, out.data()),
               "LDETable::extend_table_with_randomizers");
    return to_rowmajor(out, m, w);
}

std::pair<std::vector<Digest>, Digest> LDETable::compute_merkle_tree(
    const std::vector<std::vector<BFieldElement>>& lde_table) {
    if (lde_table.empty()) return {{}, Digest()};
    std::vector<Digest> rows = Tip5::hash_rows(lde_table);
    MerkleTree t(rows);
    return {rows, t.root()};
}

Digest LDETable::compute_lde_and_merkle(const std::vector<std::vector<BFieldElement>>& table,
                                        const ArithmeticDomain& td, const ArithmeticDomain& qd,
                                        const std::vector<std::vector<BFieldElement>>& rnd) {
    if (table.empty()) return Digest();
    size_t n = table.size(), w = table[0].size();
    if (rnd.size() != w)
        throw std::invalid_argument("Randomizer coefficients count must match column count");
    size_t k = 0;
    std::vector<uint64_t> r = bfe_rand_rows(rnd, w, &k);
    std::vector<uint64_t> cm = to_colmajor(table);
    uint64_t root[5];
    amd::ensure_device();
    amd::check(tvh_commit_host(cm.data(), w, n, k ? r.data() : nullptr, k, td.offset.value(),
                               qd.offset.value(), qd.length, nullptr, nullptr, root),
               "LDETable::compute_lde_and_merkle");
    return Digest(BFieldElement(root[0]), BFieldElement(root[1]), BFieldElement(root[2]),
                  BFieldElement(root[3]), BFieldElement(root[4]));
}

// ------------------------------------------------------------------ Tip5
namespace {
std::array<uint8_t, 256> make_lookup_table() {
    std::array<uint8_t, 256> t{};
    for (unsigned x = 0; x < 256; x++) {
        unsigned y = x + 1;
        t[x] = (uint8_t)(((y * y % 257) * y % 257) - 1);
    }
    return t;
}
std::array<int64_t, Tip5::STATE_SIZE> make_mds_column() {
    std::array<int64_t, Tip5::STATE_SIZE> c{};
    for (size_t i = 0; i < Tip5::STATE_SIZE; i++) c[i] = tip5c::kMdsFirstColumn[i];
    return c;
}
std::array<BFieldElement, Tip5::NUM_ROUNDS * Tip5::STATE_SIZE> make_round_constants() {
    std::array<BFieldElement, Tip5::NUM_ROUNDS * Tip5::STATE_SIZE> c{};
    for (size_t i = 0; i < c.size(); i++) c[i] = BFieldElement(tip5c::kRoundConstants[i]);
    return c;
}
}  // namespace

const std::array<uint8_t, 256> Tip5::LOOKUP_TABLE = make_lookup_table();
const std::array<int64_t, Tip5::STATE_SIZE> Tip5::MDS_MATRIX_FIRST_COLUMN = make_mds_column();
const std::array<BFieldElement, Tip5::NUM_ROUNDS * Tip5::STATE_SIZE> Tip5::ROUND_CONSTANTS =
    make_round_constants();

Tip5::Tip5() { state.fill(BFieldElement::zero()); }

Tip5 Tip5::init() { return Tip5(); }

void Tip5::round(size_t round_index) {
    if (round_index >= NUM_ROUNDS) throw std::out_of_range("Tip5 round index out of range");
    amd::ensure_device();
    amd::check(tvh_tip5_rounds_host(reinterpret_cast<uint64_t*>(state.data()), 1,
                                    (uint32_t)round_index, 1, nullptr),
               "Tip5::round");
}

void Tip5::permutation() {
    amd::ensure_device();
    amd::check(tvh_tip5_permutation_batch_host(reinterpret_cast<uint64_t*>(state.data()), 1),
               "Tip5::permutation");
}

void Tip5::permutation_batch(std::vector<std::array<BFieldElement, STATE_SIZE>>& states) {
    if (states.empty()) return;
    amd::ensure_device();
    amd::check(tvh_tip5_permutation_batch_host(reinterpret_cast<uint64_t*>(states.data()),
                                               states.size()),
               "Tip5::permutation_batch");
}

std::vector<std::array<BFieldElement, Tip5::STATE_SIZE>> Tip5::trace() {
    std::vector<std::array<BFieldElement, STATE_SIZE>> tr(NUM_ROUNDS + 1);
    amd::ensure_device();
    amd::check(tvh_tip5_rounds_host(reinterpret_cast<uint64_t*>(state.data()), 1, 0,
                                    (uint32_t)NUM_ROUNDS, reinterpret_cast<uint64_t*>(tr.data())),
               "Tip5::trace");
    return tr;
}

void Tip5::absorb(const std::vector<BFieldElement>& elements) {
    for (size_t i = 0; i < elements.size() && i < RATE; i++) state[i] = elements[i];
    permutation();
}

std::vector<BFieldElement> Tip5::squeeze(size_t count) {
    std::vector<BFieldElement> r;
    for (size_t i = 0; i < count && i < RATE; i++) r.push_back(state[i]);
    return r;
}

Digest Tip5::hash_10(const std::array<BFieldElement, RATE>& input) {
    Tip5 t;
    for (size_t i = 0; i < RATE; i++) t.state[i] = input[i];
    for (size_t i = RATE; i < STATE_SIZE; i++) t.state[i] = BFieldElement::one();
    t.permutation();
    return Digest(t.state[0], t.state[1], t.state[2], t.state[3], t.state[4]);
}

Digest Tip5::hash_varlen(const std::vector<BFieldElement>& input) {
    amd::ensure_device();
    uint64_t d[5];
    uint64_t dummy = 0;
    amd::check(tvh_hash_bfield_rows_host(input.empty() ? &dummy : raw(input), 1, input.size(), d),
               "Tip5::hash_varlen");
    return Digest(BFieldElement(d[0]), BFieldElement(d[1]), BFieldElement(d[2]),
                  BFieldElement(d[3]), BFieldElement(d[4]));
}

Digest Tip5::hash_pair(const Digest& l, const Digest& r) {
    return hash_pairs({l}, {r})[0];
}

std::vector<Digest> Tip5::hash_pairs(const std::vector<Digest>& l, const std::vector<Digest>& r) {
    if (l.size() != r.size()) throw std::invalid_argument("hash_pairs: length mismatch");
    std::vector<Digest> out(l.size());
    if (l.empty()) return out;
    amd::ensure_device();
    amd::check(tvh_hash_pairs_host(reinterpret_cast<const uint64_t*>(l.data()),
                                   reinterpret_cast<const uint64_t*>(r.data()),
                                   reinterpret_cast<uint64_t*>(out.data()), l.size()),
               "Tip5::hash_pairs");
    return out;
}

std::vector<Digest> Tip5::hash_rows(const std::vector<std::vector<BFieldElement>>& rows) {
    std::vector<Digest> out(rows.size());
    if (rows.empty()) return out;
    size_t w = rows[0].size();
    std::vector<uint64_t> cm = to_colmajor(rows);
    uint64_t dummy = 0;
    amd::ensure_device();
    amd::check(tvh_hash_bfield_rows_host(w ? cm.data() : &dummy, rows.size(), w,
                                         reinterpret_cast<uint64_t*>(out.data())),
               "Tip5::hash_rows");
    return out;
}

// ------------------------------------------------------------------ Merkle
MerkleTree::MerkleTree(const std::vector<Digest>& leaves) {
    if (leaves.empty()) throw std::invalid_argument("Cannot create Merkle tree with no leaves");
    if (!pow2(leaves.size())) throw std::invalid_argument("Number of leaves must be a power of 2");
    num_leaves_ = leaves.size();
    nodes_.resize(2 * num_leaves_);
    amd::ensure_device();
    amd::check(tvh_merkle_tree_host(reinterpret_cast<const uint64_t*>(leaves.data()), num_leaves_,
                                    TVH_TREE_HEAP, reinterpret_cast<uint64_t*>(nodes_.data()),
                                    nullptr),
               "MerkleTree");
    nodes_[0] = Digest();
}

MerkleTree MerkleTree::parallel_new(const std::vector<Digest>& leaves) { return MerkleTree(leaves); }

MerkleTree MerkleTree::from_heap_nodes(std::vector<Digest> nodes) {
    size_t n = nodes.size() / 2;
    if (nodes.size() % 2 || !pow2(n)) throw std::invalid_argument("heap nodes: size must be 2n, n a power of 2");
    MerkleTree t;
    t.nodes_ = std::move(nodes);
    t.nodes_[0] = Digest();
    t.num_leaves_ = n;
    return t;
}

Digest MerkleTree::root() const { return nodes_[1]; }

size_t MerkleTree::height() const {
    size_t h = 0;
    while ((size_t(1) << h) < num_leaves_) h++;
    return h;
}

Digest MerkleTree::leaf(size_t i) const {
    if (i >= num_leaves_) throw std::out_of_range("Leaf index out of range");
    return nodes_[num_leaves_ + i];
}

std::vector<Digest> MerkleTree::authentication_path(size_t leaf_index) const {
    if (leaf_index >= num_leaves_) throw std::out_of_range("Leaf index out of range");
    size_t cnt = 0;
    amd::check(tvh_auth_path_indices(num_leaves_, leaf_index, nullptr, 0, &cnt), "auth path");
    std::vector<uint64_t> idx(cnt);
    amd::check(tvh_auth_path_indices(num_leaves_, leaf_index, idx.data(), cnt, &cnt), "auth path");
    std::vector<Digest> p;
    for (uint64_t v : idx) p.push_back(nodes_[v]);
    return p;
}

std::vector<Digest> MerkleTree::authentication_structure(const std::vector<size_t>& li) const {
    size_t cnt = 0;
    amd::check(tvh_auth_structure_indices(num_leaves_, li.data(), li.size(), nullptr, 0, &cnt),
               "MerkleTree::authentication_structure");
    std::vector<uint64_t> idx(cnt);
    amd::check(tvh_auth_structure_indices(num_leaves_, li.data(), li.size(), idx.data(), cnt, &cnt),
               "MerkleTree::authentication_structure");
    std::vector<Digest> s;
    for (uint64_t v : idx) s.push_back(nodes_[v]);
    return s;
}

bool MerkleTree::verify(const Digest& root, size_t index, const Digest& leaf,
                        const std::vector<Digest>& path) {
    Digest cur = leaf;
    for (const Digest& sib : path) {
        cur = (index % 2 == 0) ? Tip5::hash_pair(cur, sib) : Tip5::hash_pair(sib, cur);
        index /= 2;
    }
    return cur == root;
}

bool MerkleTree::verify_authentication_structure(const Digest& root, size_t n,
                                                 const std::vector<size_t>& li,
                                                 const std::vector<Digest>& leaves,
                                                 const std::vector<Digest>& structure) {
    // merkle_tree.cpp:267-341: rebuild the root from leaves + structure
    if (li.size() != leaves.size()) return false;
    if (!pow2(n)) return false;
    if (n == 1) return li.size() == 1 && li[0] == 0 && leaves[0] == root;
    size_t cnt = 0;
    if (tvh_auth_structure_indices(n, li.data(), li.size(), nullptr, 0, &cnt)) return false;
    std::vector<uint64_t> idx(cnt);
    if (tvh_auth_structure_indices(n, li.data(), li.size(), idx.data(), cnt, &cnt)) return false;
    if (idx.size() != structure.size()) return false;
    std::unordered_map<size_t, Digest> known;
    for (size_t i = 0; i < li.size(); i++) known[n + li[i]] = leaves[i];
    for (size_t i = 0; i < idx.size(); i++) known[idx[i]] = structure[i];
    // hash level by level, all pairs of a level in one batch
    for (size_t level_lo = n; level_lo > 1; level_lo /= 2) {
        std::vector<size_t> parents;
        std::vector<Digest> l, r;
        for (const auto& kv : known) {
            size_t v = kv.first;
            if (v < level_lo || v >= 2 * level_lo || (v & 1)) continue;
            auto it = known.find(v + 1);
            if (it == known.end() || known.count(v / 2)) continue;
            parents.push_back(v / 2);
            l.push_back(kv.second);
            r.push_back(it->second);
        }
        if (parents.empty()) continue;
        std::vector<Digest> h = Tip5::hash_pairs(l, r);
        for (size_t i = 0; i < parents.size(); i++) known[parents[i]] = h[i];
    }
    auto it = known.find(1);
    return it != known.end() && it->second == root;
}

// ------------------------------------------------------------------ TableCommitment
TableCommitment::TableCommitment(std::unique_ptr<MerkleTree> tree) : tree_(std::move(tree)) {}

TableCommitment TableCommitment::from_digests(const std::vector<Digest>& row_digests) {
    return TableCommitment(std::make_unique<MerkleTree>(row_digests));
}

namespace {
// Row digests of a column-major table, then the heap-layout tree, in one GPU
// round trip (tvh_hash_bfield_rows_host, tvh_merkle_tree_host).
std::unique_ptr<MerkleTree> commit_colmajor(const std::vector<uint64_t>& cm, size_t rows,
                                            size_t width) {
    if (rows == 0) throw std::invalid_argument("Cannot create Merkle tree with no leaves");
    if (!pow2(rows)) throw std::invalid_argument("Number of leaves must be a power of 2");
    amd::ensure_device();
    std::vector<Digest> dig(rows);
    uint64_t dummy = 0;
    amd::check(tvh_hash_bfield_rows_host(width ? cm.data() : &dummy, rows, width,
                                         reinterpret_cast<uint64_t*>(dig.data())),
               "TableCommitment::commit");
    std::vector<Digest> nodes(2 * rows);
    amd::check(tvh_merkle_tree_host(reinterpret_cast<const uint64_t*>(dig.data()), rows,
                                    TVH_TREE_HEAP, reinterpret_cast<uint64_t*>(nodes.data()), nullptr),
               "TableCommitment::commit");
    return std::make_unique<MerkleTree>(MerkleTree::from_heap_nodes(std::move(nodes)));
}
}  // namespace

TableCommitment TableCommitment::commit_bfield_rows(const BFieldElement* const* rows, size_t n,
                                                    size_t w) {
    std::vector<uint64_t> cm(n * w);
    for (size_t r = 0; r < n; r++)
        for (size_t c = 0; c < w; c++) cm[c * n + r] = rows[r][c].value();
    return TableCommitment(commit_colmajor(cm, n, w));
}

TableCommitment TableCommitment::commit_xfield_rows(const XFieldElement* const* rows, size_t n,
                                                    size_t w) {
    // flattened (c0, c1, c2) per element: hash_xfield_row's order
    std::vector<uint64_t> cm(n * 3 * w);
    for (size_t r = 0; r < n; r++)
        for (size_t c = 0; c < w; c++)
            for (size_t k = 0; k < 3; k++) cm[(3 * c + k) * n + r] = rows[r][c].coeff(k).value();
    return TableCommitment(commit_colmajor(cm, n, 3 * w));
}

TableCommitment TableCommitment::commit_bfield_rows(const std::vector<std::vector<BFieldElement>>& rows) {
    size_t w = rows.empty() ? 0 : rows[0].size();
    std::vector<const BFieldElement*> p(rows.size());
    for (size_t i = 0; i < rows.size(); i++) {
        if (rows[i].size() != w) throw std::invalid_argument("TableCommitment: ragged table");
        p[i] = rows[i].data();
    }
    return commit_bfield_rows(p.data(), p.size(), w);
}

TableCommitment TableCommitment::commit_xfield_rows(const std::vector<std::vector<XFieldElement>>& rows) {
    size_t w = rows.empty() ? 0 : rows[0].size();
    std::vector<const XFieldElement*> p(rows.size());
    for (size_t i = 0; i < rows.size(); i++) {
        if (rows[i].size() != w) throw std::invalid_argument("TableCommitment: ragged table");
        p[i] = rows[i].data();
    }
    return commit_xfield_rows(p.data(), p.size(), w);
}

Digest TableCommitment::root() const { return tree_->root(); }
size_t TableCommitment::num_rows() const { return tree_->num_leaves(); }
std::vector<Digest> TableCommitment::authentication_path(size_t row_index) const {
    return tree_->authentication_path(row_index);
}
std::vector<Digest> TableCommitment::authentication_structure(const std::vector<size_t>& idx) const {
    return tree_->authentication_structure(idx);
}
bool TableCommitment::verify(const Digest& root, size_t row_index, const Digest& row_digest,
                             const std::vector<Digest>& auth_path) {
    return MerkleTree::verify(root, row_index, row_digest, auth_path);
}

Digest hash_row(const std::vector<BFieldElement>& row) {
    if (row.empty()) return Digest::zero();
    return Tip5::hash_varlen(row);
}

Digest hash_bfield_row(const std::vector<BFieldElement>& row) { return Tip5::hash_varlen(row); }

Digest hash_xfield_row(const std::vector<XFieldElement>& row) {
    std::vector<BFieldElement> flat;
    flat.reserve(3 * row.size());
    for (const auto& x : row)
        for (size_t j = 0; j < 3; j++) flat.push_back(x.coeff(j));
    return Tip5::hash_varlen(flat);
}

}  // namespace triton_vm

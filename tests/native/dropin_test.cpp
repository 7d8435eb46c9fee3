// Exercises the drop-in C++ classes (the reference's include paths:
// ntt/ntt.hpp, lde/*.hpp, hash/tip5.hpp, merkle/merkle_tree.hpp,
// table/table_commitment.hpp, types/*.hpp) on the GPU the way the reference's
// own unit tests exercise its classes.  Prints one JSON object per line;
// tests/test_gpu_dropin.py checks them against the golden fixtures and the
// oracle.  Exit status is non-zero on an unexpected exception.
#include <cstdio>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "hash/tip5.hpp"
#include "lde/lde_randomized.hpp"
#include "lde/lde_table.hpp"
#include "merkle/merkle_tree.hpp"
#include "ntt/ntt.hpp"
#include "table/table_commitment.hpp"
#include "types/b_field_element.hpp"
#include "types/digest.hpp"
#include "types/x_field_element.hpp"

using namespace triton_vm;

static std::vector<BFieldElement> splitmix(uint64_t seed, size_t n) {
    std::vector<BFieldElement> out;
    uint64_t s = seed;
    while (out.size() < n) {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        if (z < BFieldElement::MODULUS) out.push_back(BFieldElement(z));
    }
    return out;
}

static void print_vec(const char* key, const std::vector<BFieldElement>& v) {
    std::printf("{\"%s\": [", key);
    for (size_t i = 0; i < v.size(); i++) std::printf("%s%llu", i ? "," : "", (unsigned long long)v[i].value());
    std::printf("]}\n");
}

static void print_digests(const char* key, const std::vector<Digest>& d) {
    std::vector<BFieldElement> flat;
    for (const auto& x : d)
        for (size_t i = 0; i < 5; i++) flat.push_back(x[i]);
    print_vec(key, flat);
}

template <class E, class F>
static void expect_throw(const char* key, F f) {
    bool ok = false;
    try {
        f();
    } catch (const E&) {
        ok = true;
    } catch (...) {
    }
    std::printf("{\"%s\": %s}\n", key, ok ? "true" : "false");
}

int main() {
    try {
        // NTT: Rust KAT [1,4,0,0] and a 64-point vector
        std::vector<BFieldElement> kat = {BFieldElement(1), BFieldElement(4), BFieldElement(0), BFieldElement(0)};
        NTT::forward(kat);
        print_vec("ntt_kat", kat);
        auto x = splitmix(4006, 64);
        auto y = x;
        NTT::forward(y);
        print_vec("ntt64_fwd", y);
        NTT::inverse(y);
        std::printf("{\"ntt64_roundtrip\": %s}\n", y == x ? "true" : "false");
        expect_throw<std::invalid_argument>("ntt_non_pow2_throws", [] {
            std::vector<BFieldElement> v(3);
            NTT::forward(v);
        });

        // LDE::extend_column == NTT::evaluate_on_coset(interpolate)
        auto tc = splitmix(31, 64);
        auto lde = LDE::extend_column(tc, ArithmeticDomain::of_length(64),
                                      ArithmeticDomain::of_length(256).with_offset(BFieldElement(7)));
        auto ref = NTT::evaluate_on_coset(NTT::interpolate(tc), 256, BFieldElement(7));
        std::printf("{\"lde_matches_coset_eval\": %s}\n", lde == ref ? "true" : "false");

        // RandomizedLDE golden case (16 -> 64, 3 randomizers)
        auto rt = splitmix(5002, 16);
        auto rr = splitmix(6002, 3);
        print_vec("rlde_16_64_k3", RandomizedLDE::extend_column_with_randomizer(
                                      rt, ArithmeticDomain::of_length(16),
                                      ArithmeticDomain::of_length(64).with_offset(BFieldElement(7)), rr));
        auto zf = RandomizedLDE::compute_zerofier(ArithmeticDomain::of_length(8).with_offset(BFieldElement(3)));
        std::printf("{\"zerofier_vanishes\": %s}\n",
                    RandomizedLDE::verify_zerofier(ArithmeticDomain::of_length(8).with_offset(BFieldElement(3)), zf) ? "true" : "false");

        // XFE golden case (16 -> 64, 3 XFE randomizers, trace offset 3)
        auto xt = splitmix(7003, 48);
        auto xr = splitmix(8003, 9);
        std::vector<XFieldElement> xcol(16), xrand(3);
        for (size_t i = 0; i < 16; i++) xcol[i] = XFieldElement(xt[3 * i], xt[3 * i + 1], xt[3 * i + 2]);
        for (size_t i = 0; i < 3; i++) xrand[i] = XFieldElement(xr[3 * i], xr[3 * i + 1], xr[3 * i + 2]);
        auto xo = RandomizedLDE::extend_xfield_column_with_xfield_randomizer(
            xcol, ArithmeticDomain::of_length(16).with_offset(BFieldElement(3)),
            ArithmeticDomain::of_length(64).with_offset(BFieldElement(7)), xrand);
        std::vector<BFieldElement> xflat;
        for (const auto& e : xo)
            for (size_t j = 0; j < 3; j++) xflat.push_back(e.coeff(j));
        print_vec("xlde_16_64_k3_x", xflat);

        // Tip5 KATs (Rust hash10_test_vectors, hash_varlen_test_vectors)
        std::array<BFieldElement, 10> pre{};
        Digest d;
        for (int i = 0; i < 6; i++) {
            d = Tip5::hash_10(pre);
            for (int j = 0; j < 5; j++) pre[i + j] = d[j];
        }
        print_digests("hash10_kat", {Tip5::hash_10(pre)});
        Digest sum;
        for (int i = 0; i < 20; i++) {
            std::vector<BFieldElement> p;
            for (int j = 0; j < i; j++) p.push_back(BFieldElement(j));
            Digest h = Tip5::hash_varlen(p);
            for (int j = 0; j < 5; j++) sum[j] = sum[j] + h[j];
        }
        print_digests("hash_varlen_kat_sum", {sum});
        std::printf("{\"hash_row_empty_is_zero\": %s}\n", hash_row({}) == Digest::zero() ? "true" : "false");

        // Merkle: golden 64-leaf tree, auth structures, verification
        auto lv = splitmix(9000, 320);
        std::vector<Digest> leaves(64);
        for (size_t i = 0; i < 64; i++) leaves[i] = Digest(lv[5 * i], lv[5 * i + 1], lv[5 * i + 2], lv[5 * i + 3], lv[5 * i + 4]);
        MerkleTree t(leaves);
        print_digests("merkle64_root", {t.root()});
        print_digests("auth_5_9_40", t.authentication_structure({5, 9, 40}));
        auto path = t.authentication_path(17);
        std::printf("{\"verify_ok\": %s, \"verify_wrong_index\": %s}\n",
                    MerkleTree::verify(t.root(), 17, t.leaf(17), path) ? "true" : "false",
                    MerkleTree::verify(t.root(), 16, t.leaf(17), path) ? "true" : "false");
        std::vector<size_t> qi = {5, 9, 40};
        std::vector<Digest> ql = {t.leaf(5), t.leaf(9), t.leaf(40)};
        std::printf("{\"verify_structure_ok\": %s}\n",
                    MerkleTree::verify_authentication_structure(t.root(), 64, qi, ql, t.authentication_structure(qi)) ? "true" : "false");
        expect_throw<std::invalid_argument>("merkle_non_pow2_throws", [] { MerkleTree m(std::vector<Digest>(3)); });
        expect_throw<std::out_of_range>("merkle_leaf_range_throws", [&] { t.leaf(64); });

        // Tip5 public tables, per-round states (Tip5::trace / Tip5::round)
        {
            std::vector<BFieldElement> lut, mds;
            for (uint8_t v : Tip5::LOOKUP_TABLE) lut.push_back(BFieldElement(v));
            for (int64_t v : Tip5::MDS_MATRIX_FIRST_COLUMN) mds.push_back(BFieldElement((uint64_t)v));
            print_vec("tip5_lookup_table", lut);
            print_vec("tip5_mds_first_column", mds);
            print_vec("tip5_round_constants",
                      std::vector<BFieldElement>(Tip5::ROUND_CONSTANTS.begin(), Tip5::ROUND_CONSTANTS.end()));
            auto s0 = splitmix(77, 16);
            Tip5 a;
            for (size_t i = 0; i < 16; i++) a.state[i] = s0[i];
            Tip5 b = a, c = a;
            auto tr = a.trace();
            std::vector<BFieldElement> flat;
            for (const auto& st : tr) flat.insert(flat.end(), st.begin(), st.end());
            print_vec("tip5_trace_77", flat);
            bool rounds_match = true;
            for (size_t r = 0; r < Tip5::NUM_ROUNDS; r++) {
                b.round(r);
                rounds_match = rounds_match && b.state == tr[r + 1];
            }
            c.permutation();
            std::printf("{\"tip5_rounds_match_trace\": %s, \"tip5_trace_leaves_permuted\": %s}\n",
                        rounds_match ? "true" : "false", (a.state == c.state && c.state == tr.back()) ? "true" : "false");
            expect_throw<std::out_of_range>("tip5_round_range_throws", [&] { b.round(5); });
        }

        // Value types: field inverses, Montgomery import, hex round trip
        {
            auto v = splitmix(99, 30);
            bool inv_ok = true;
            std::vector<XFieldElement> xs;
            for (size_t i = 0; i + 2 < v.size(); i += 3) {
                XFieldElement x(v[i], v[i + 1], v[i + 2]);
                xs.push_back(x);
                inv_ok = inv_ok && (x * x.inverse()).is_one() && (x / x).is_one() &&
                         x.pow(5) * x.pow(7) == x.pow(12);
            }
            auto xinv = XFieldElement::batch_inversion(xs);
            for (size_t i = 0; i < xs.size(); i++) inv_ok = inv_ok && xinv[i] == xs[i].inverse();
            auto binv = BFieldElement::batch_inversion(v);
            for (size_t i = 0; i < v.size(); i++) inv_ok = inv_ok && (binv[i] * v[i]).is_one();
            // Montgomery form of x is x * 2^64 mod p = x * (2^32 - 1)
            bool raw_ok = true;
            for (const auto& x : v)
                raw_ok = raw_ok && BFieldElement::from_raw_u64((x * BFieldElement(BFieldElement::MONTGOMERY_R)).value()) == x;
            Digest dg(v[0], v[1], v[2], v[3], v[4]);
            std::printf("{\"field_inverses_ok\": %s, \"from_raw_u64_ok\": %s, \"hex_roundtrip\": %s, \"xfe_string\": \"%s\"}\n",
                        inv_ok ? "true" : "false", raw_ok ? "true" : "false",
                        Digest::from_hex(dg.to_hex()) == dg ? "true" : "false",
                        XFieldElement(BFieldElement(1), BFieldElement(2), BFieldElement(3)).to_string().c_str());
            expect_throw<std::domain_error>("xfe_zero_inverse_throws", [] { XFieldElement::zero().inverse(); });
        }

        // NTT::interpolate from 8 host threads at once (each call uses its own
        // device buffers): every result must equal its single-threaded value
        {
            const int T = 8;
            std::vector<std::vector<BFieldElement>> cols(T), got(T);
            for (int t = 0; t < T; t++) cols[t] = splitmix(500 + t, size_t(1) << (8 + t % 5));
            std::vector<std::thread> th;
            bool ok = true;
            std::vector<int> same(T, 1);
            for (int t = 0; t < T; t++)
                th.emplace_back([&, t] {
                    got[t] = NTT::interpolate(cols[t]);
                    for (int rep = 0; rep < 6; rep++)
                        if (NTT::interpolate(cols[t]) != got[t]) same[t] = 0;
                });
            for (auto& x : th) x.join();
            for (int t = 0; t < T; t++) ok = ok && same[t];
            std::printf("{\"threaded_interpolate_stable\": %s}\n", ok ? "true" : "false");
            for (int t = 0; t < T; t++) {
                std::string key = "threaded_interpolate_" + std::to_string(t);
                print_vec(key.c_str(), got[t]);
            }
        }

        // TableCommitment over an extended 2^10 x 379 table, and from digests
        {
            size_t n = 1024, w = 379;
            auto flat = splitmix(1, n * w);
            std::vector<std::vector<BFieldElement>> table(n, std::vector<BFieldElement>(w));
            for (size_t r = 0; r < n; r++)
                for (size_t c = 0; c < w; c++) table[r][c] = flat[r * w + c];
            auto ext = LDETable::extend_table_with_randomizers(
                table, ArithmeticDomain::of_length(n),
                ArithmeticDomain::of_length(4 * n).with_offset(BFieldElement(7)),
                std::vector<std::vector<BFieldElement>>(w));
            TableCommitment tc1 = TableCommitment::commit_bfield_rows(ext);
            print_digests("table_commitment_1024_379", {tc1.root()});
            std::vector<Digest> rows(ext.size());
            for (size_t r = 0; r < ext.size(); r++) rows[r] = hash_bfield_row(ext[r]);
            TableCommitment tc2 = TableCommitment::from_digests(rows);
            auto p = tc2.authentication_path(1234);
            std::printf("{\"table_commitment_from_digests_matches\": %s, \"table_commitment_rows\": %zu, "
                        "\"table_commitment_verify\": %s, \"table_commitment_tree_root\": %s}\n",
                        tc2.root() == tc1.root() ? "true" : "false", tc1.num_rows(),
                        TableCommitment::verify(tc1.root(), 1234, rows[1234], p) ? "true" : "false",
                        tc1.tree().root() == tc1.root() ? "true" : "false");
            auto st = tc1.authentication_structure({3, 77, 4000});
            print_digests("table_commitment_auth_3_77_4000", st);
            // XFE rows: hash_xfield_row flattening, 64 rows x 5 XFE columns
            auto xv = splitmix(321, 64 * 15);
            std::vector<std::vector<XFieldElement>> xrows(64, std::vector<XFieldElement>(5));
            for (size_t r = 0; r < 64; r++)
                for (size_t c = 0; c < 5; c++)
                    xrows[r][c] = XFieldElement(xv[15 * r + 3 * c], xv[15 * r + 3 * c + 1], xv[15 * r + 3 * c + 2]);
            TableCommitment tx = TableCommitment::commit_xfield_rows(xrows);
            std::vector<Digest> xd(64);
            for (size_t r = 0; r < 64; r++) xd[r] = hash_xfield_row(xrows[r]);
            std::printf("{\"xfield_commitment_matches\": %s}\n",
                        TableCommitment::from_digests(xd).root() == tx.root() ? "true" : "false");
            print_digests("xfield_commitment_root", {tx.root()});
            expect_throw<std::invalid_argument>("table_commitment_empty_throws",
                                                [] { TableCommitment::from_digests({}); });
        }

        // Whole commits (LDETable::compute_lde_and_merkle): golden shapes
        const size_t shapes[][3] = {{8, 7, 1}, {16, 9, 12}, {1024, 379, 1}};
        for (const auto& sh : shapes) {
            size_t n = sh[0], w = sh[1];
            auto flat = splitmix(sh[2], n * w);  // row-major draw order
            std::vector<std::vector<BFieldElement>> table(n, std::vector<BFieldElement>(w));
            for (size_t r = 0; r < n; r++)
                for (size_t c = 0; c < w; c++) table[r][c] = flat[r * w + c];
            Digest root = LDETable::compute_lde_and_merkle(
                table, ArithmeticDomain::of_length(n),
                ArithmeticDomain::of_length(4 * n).with_offset(BFieldElement(7)),
                std::vector<std::vector<BFieldElement>>(w));
            std::string key = "commit_" + std::to_string(n) + "_" + std::to_string(w);
            print_digests(key.c_str(), {root});
            if (n == 8) {
                auto ext = LDETable::extend_table_with_randomizers(
                    table, ArithmeticDomain::of_length(n),
                    ArithmeticDomain::of_length(4 * n).with_offset(BFieldElement(7)),
                    std::vector<std::vector<BFieldElement>>(w));
                auto [rows, r2] = LDETable::compute_merkle_tree(ext);
                std::printf("{\"two_step_commit_matches\": %s, \"cw00\": %llu}\n", r2 == root ? "true" : "false",
                            (unsigned long long)ext[0][0].value());
            }
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "unexpected exception: %s\n", e.what());
        return 1;
    }
    return 0;
}

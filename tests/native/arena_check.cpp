// Randomised check of the commit server's first-fit arena (csrc/arena.hpp)
// against a brute-force model: a granule bitmap searched lowest-offset-first.
// Prints one JSON line; exits non-zero on the first disagreement.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "arena.hpp"

int main() {
    const size_t G = 64;  // granules of the model
    std::mt19937_64 rng(12345);
    long checks = 0;
    for (int trial = 0; trial < 200; trial++) {
        tvh::FirstFitArena a;
        a.reset(G);
        std::vector<int> owner(G, -1);
        std::vector<std::pair<size_t, size_t>> live;  // (off, len)
        for (int step = 0; step < 400; step++) {
            bool do_alloc = live.empty() || (rng() % 3 != 0);
            if (do_alloc) {
                size_t len = 1 + rng() % 12;
                // model: lowest offset with len free granules
                size_t want = SIZE_MAX;
                for (size_t o = 0; o + len <= G && want == SIZE_MAX; o++) {
                    bool ok = true;
                    for (size_t k = 0; k < len && ok; k++) ok = owner[o + k] < 0;
                    if (ok) want = o;
                }
                // first-fit over holes sorted by offset finds the first HOLE that is
                // large enough, which is the lowest feasible offset
                size_t off = 0;
                bool got = a.alloc(len, &off);
                if (got != (want != SIZE_MAX) || (got && off != want)) {
                    std::printf("{\"ok\": false, \"trial\": %d, \"step\": %d, \"len\": %zu}\n", trial,
                                step, len);
                    return 1;
                }
                if (got) {
                    for (size_t k = 0; k < len; k++) owner[off + k] = (int)live.size();
                    live.push_back({off, len});
                }
            } else {
                size_t i = rng() % live.size();
                auto [off, len] = live[i];
                a.release(off, len);
                for (size_t k = 0; k < len; k++) owner[off + k] = -1;
                live.erase(live.begin() + i);
                for (size_t k = 0; k < G; k++)
                    if (owner[k] >= (int)i) owner[k]--;
            }
            size_t used = 0, holes = 0, run = 0, best = 0;
            for (size_t k = 0; k < G; k++) {
                used += owner[k] >= 0;
                if (owner[k] < 0) {
                    if (k == 0 || owner[k - 1] >= 0) holes++;
                    run++;
                    best = run > best ? run : best;
                } else {
                    run = 0;
                }
            }
            if (a.used() != used || a.holes() != holes || a.largest_hole() != best) {
                std::printf("{\"ok\": false, \"trial\": %d, \"step\": %d, \"used\": %zu, \"model\": %zu}\n",
                            trial, step, a.used(), used);
                return 1;
            }
            checks++;
        }
        for (auto [off, len] : live) a.release(off, len);
        if (a.used() != 0 || a.holes() != 1 || a.largest_hole() != G) {
            std::printf("{\"ok\": false, \"trial\": %d, \"final\": true}\n", trial);
            return 1;
        }
    }
    std::printf("{\"ok\": true, \"checks\": %ld}\n", checks);
    return 0;
}

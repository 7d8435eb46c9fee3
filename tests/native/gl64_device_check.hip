// Device check of the Goldilocks primitives in triton-vm-prover_amd/csrc/gl64.hpp
// as the gfx950 compiler builds them, against unsigned __int128 arithmetic on
// the host.  The host check (gl64_check.cpp) runs the same source through the
// host compiler; this one catches code-generation differences of the carry
// chains on the device.  Inputs: edge values around p and 2^32 crossed with
// each other, then random operands.  Every kernel is bounds-checked against n.
#include "gl64.hpp"

#include <cstdio>
#include <random>
#include <vector>

typedef unsigned __int128 u128;
using gl::u32;
using gl::u64;

enum Op {
    OP_ADD, OP_SUB, OP_CANON, OP_MONT_MUL, OP_MONT_SQR, OP_MONTYRED, OP_TO_MONT, OP_ADD_LAZY,
    OP_MUL_LO, OP_MUL_HI, OP_SQR_LO, OP_SQR_HI, OP_LIN52, OP_LIN52_CANON, OP_MUL,
    OP_TO_MONT_CANON, OP_MONT_MUL_CANON, OP_ADD_LAZY_CANON, OP_MONT_SQR_CANON,
    OP_MONTYRED_CANON, OP_SUB_CANON, OP_ADD_CANON, OP_POW2_ADD, OP_COUNT
};
static const char* kNames[OP_COUNT] = {"add", "sub", "canon", "mont_mul", "mont_sqr", "montyred",
                                       "to_mont", "add_lazy", "mul_wide lo", "mul_wide hi",
                                       "sqr_wide lo", "sqr_wide hi", "lin52", "canon(lin52)",
                                       "mul", "canon(to_mont)", "canon(mont_mul)",
                                       "canon(add_lazy)", "canon(mont_sqr)", "canon(montyred)",
                                       "canon(sub)", "canon(add)", "add(sub, 2^40 x)"};

__global__ void k_ops(int op, const u64* __restrict__ a, const u64* __restrict__ b,
                      u64* __restrict__ out, u32 n) {
    u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u64 x = a[i], y = b[i], lo, hi;
    u64 r = 0;
    switch (op) {
        case OP_ADD: r = gl::add(x % gl::P, y % gl::P); break;
        case OP_SUB: r = gl::sub(x % gl::P, y % gl::P); break;
        case OP_CANON: r = gl::canon(x); break;
        case OP_MONT_MUL: r = gl::mont_mul(x, y); break;
        case OP_MONT_SQR: r = gl::mont_sqr(x); break;
        case OP_MONTYRED: r = gl::montyred(x, y % gl::P); break;
        case OP_TO_MONT: r = gl::to_mont(x); break;
        case OP_ADD_LAZY: r = gl::add_lazy(x, y % gl::P); break;
        case OP_MUL_LO: gl::mul_wide_4mad(x, y, lo, hi); r = lo; break;
        case OP_MUL_HI: gl::mul_wide_4mad(x, y, lo, hi); r = hi; break;
        case OP_SQR_LO: gl::sqr_wide(x, lo, hi); r = lo; break;
        case OP_SQR_HI: gl::sqr_wide(x, lo, hi); r = hi; break;
        case OP_LIN52: r = gl::lin52(x & ((1ull << 52) - 1), y & ((1ull << 52) - 1)); break;
        case OP_LIN52_CANON:
            r = gl::canon(gl::lin52(x & ((1ull << 52) - 1), y & ((1ull << 52) - 1)));
            break;
        case OP_MUL: r = gl::mul(x, y); break;
        case OP_TO_MONT_CANON: r = gl::canon(gl::to_mont(x)); break;
        case OP_MONT_MUL_CANON: r = gl::canon(gl::mont_mul(x, y)); break;
        case OP_ADD_LAZY_CANON: r = gl::canon(gl::add_lazy(x, y % gl::P)); break;
        case OP_MONT_SQR_CANON: r = gl::canon(gl::mont_sqr(x)); break;
        case OP_MONTYRED_CANON: r = gl::canon(gl::montyred(x, y)); break;
        case OP_SUB_CANON: r = gl::canon(gl::sub(x % gl::P, y % gl::P)); break;
        case OP_ADD_CANON: r = gl::canon(gl::add(x % gl::P, y % gl::P)); break;
        case OP_POW2_ADD: {  // a butterfly chain: sub, then a shift multiply, then add
            u64 d = gl::sub(x % gl::P, y % gl::P);
            r = gl::add(d, gl::mul_pow2<40>(d));
            break;
        }
        default: break;
    }
    out[i] = r;
}

template <int E>
__global__ void k_pow2(const u64* __restrict__ a, u64* __restrict__ out, u32 n) {
    u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = gl::mul_pow2<E>(a[i]);
}

static u64 pmod(u128 v) { return (u64)(v % gl::P); }
static const u64 RINV = gl::inv(0xFFFFFFFFull);  // 2^-64 mod p
static u64 pow2_ref(u64 x, int e) { return pmod((u128)(x % gl::P) * gl::pow2_mod_p(e)); }

// expected value, or the residue class for "lazy" results (checked mod p)
static bool check(int op, u64 x, u64 y, u64 got) {
    const u64 P = gl::P;
    u128 w = (u128)x * y;
    switch (op) {
        case OP_ADD: return got == pmod((u128)(x % P) + (y % P));
        case OP_SUB: return got == pmod((u128)(x % P) + P - (y % P));
        case OP_CANON: return got == x % P;
        case OP_MONT_MUL: {
            u64 want = pmod((u128)pmod(w) * RINV);
            if (got % P != want) return false;
            return (u64)(w >> 64) >= P || got < P;  // canonical when hi < p
        }
        case OP_MONT_SQR: return got % P == pmod((u128)pmod((u128)x * x) * RINV);
        case OP_MONTYRED: {
            u128 v = ((u128)(y % P) << 64) | x;
            return got % P == pmod((u128)pmod(v) * RINV) && got < P;
        }
        case OP_TO_MONT: return got % P == pmod((u128)x * 0xFFFFFFFFull);
        case OP_ADD_LAZY: return got % P == pmod((u128)x + (y % P));
        case OP_MUL_LO: return got == (u64)w;
        case OP_MUL_HI: return got == (u64)(w >> 64);
        case OP_SQR_LO: return got == (u64)((u128)x * x);
        case OP_SQR_HI: return got == (u64)(((u128)x * x) >> 64);
        case OP_LIN52:
        case OP_LIN52_CANON: {
            const u64 M52 = (1ull << 52) - 1;
            u64 want = pmod((u128)(x & M52) + ((u128)(y & M52) << 32));
            return op == OP_LIN52 ? got % P == want : got == want;
        }
        case OP_MUL: return got == pmod(w);
        case OP_TO_MONT_CANON: return got == pmod((u128)x * 0xFFFFFFFFull);
        case OP_MONT_MUL_CANON: return got == pmod((u128)pmod(w) * RINV);
        case OP_ADD_LAZY_CANON: return got == pmod((u128)x + (y % P));
        case OP_MONT_SQR_CANON: return got == pmod((u128)pmod((u128)x * x) * RINV);
        case OP_MONTYRED_CANON: return got == pmod((u128)pmod(((u128)y << 64) | x) * RINV);
        case OP_SUB_CANON: return got == pmod((u128)(x % P) + P - (y % P));
        case OP_ADD_CANON: return got == pmod((u128)(x % P) + (y % P));
        case OP_POW2_ADD: {
            u64 d = pmod((u128)(x % P) + P - (y % P));
            return got == pmod((u128)d + pmod((u128)d * gl::pow2_mod_p(40)));
        }
        default: return false;
    }
}

template <int E>
static int run_pow2(const std::vector<u64>& a, u64* da, u64* dout, u32 n, std::vector<u64>& host) {
    hipLaunchKernelGGL(k_pow2<E>, dim3((n + 255) / 256), dim3(256), 0, 0, da, dout, n);
    if (hipMemcpy(host.data(), dout, n * 8ull, hipMemcpyDeviceToHost) != hipSuccess) return 1 << 30;
    int bad = 0;
    for (u32 i = 0; i < n; i++)
        if (host[i] != pow2_ref(a[i], E)) {
            if (bad < 2)
                printf("mul_pow2<%d>(%016llx) = %016llx\n", E, (unsigned long long)a[i],
                       (unsigned long long)host[i]);
            bad++;
        }
    return bad;
}
template <int... Es>
static int all_pow2(std::integer_sequence<int, Es...>, const std::vector<u64>& a, u64* da,
                    u64* dout, u32 n, std::vector<u64>& host) {
    int bad = 0;
    ((bad += run_pow2<Es>(a, da, dout, n, host)), ...);
    return bad;
}

int main() {
    const u64 P = gl::P;
    const u64 edge[] = {0, 1, 2, 3, P - 1, P - 2, P, P + 1, 0xFFFFFFFFull, 0x100000000ull,
                        0x100000001ull, P - 0x100000000ull, 0xFFFFFFFFFFFFFFFFull,
                        0xFFFFFFFF00000000ull, 0x8000000000000000ull, 0x7FFFFFFFFFFFFFFFull,
                        0xFFFFFFFEFFFFFFFFull, 0x00000000FFFFFFFEull, (1ull << 52) - 1, 1ull << 52};
    const u32 ne = sizeof(edge) / sizeof(edge[0]);
    const u32 n = 1u << 20;
    std::vector<u64> a(n), b(n), host(n);
    std::mt19937_64 rng(11);
    for (u32 i = 0; i < n; i++) {
        if (i < ne * ne) {
            a[i] = edge[i % ne];
            b[i] = edge[i / ne];
        } else {
            // random words, some with a few edge bits set or small / near-p values
            a[i] = rng();
            b[i] = rng();
            if ((i & 7) == 1) a[i] = P - (rng() & 0xFFFFFFFFull);
            if ((i & 7) == 2) b[i] = rng() & 0xFFFFFFFFull;
            if ((i & 7) == 3) a[i] |= 0xFFFFFFFF00000000ull;
        }
    }
    u64 *da = nullptr, *db = nullptr, *dout = nullptr;
    if (hipMalloc(&da, n * 8ull) != hipSuccess || hipMalloc(&db, n * 8ull) != hipSuccess ||
        hipMalloc(&dout, n * 8ull) != hipSuccess) {
        printf("no device memory\n");
        return 2;
    }
    if (hipMemcpy(da, a.data(), n * 8ull, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, b.data(), n * 8ull, hipMemcpyHostToDevice) != hipSuccess) {
        printf("copy failed\n");
        return 2;
    }
    int total = 0;
    for (int op = 0; op < OP_COUNT; op++) {
        hipLaunchKernelGGL(k_ops, dim3((n + 255) / 256), dim3(256), 0, 0, op, da, db, dout, n);
        if (hipMemcpy(host.data(), dout, n * 8ull, hipMemcpyDeviceToHost) != hipSuccess) {
            printf("kernel or copy failed\n");
            return 2;
        }
        int bad = 0;
        for (u32 i = 0; i < n; i++)
            if (!check(op, a[i], b[i], host[i])) {
                if (bad < 2)
                    printf("%s(%016llx, %016llx) = %016llx\n", kNames[op],
                           (unsigned long long)a[i], (unsigned long long)b[i],
                           (unsigned long long)host[i]);
                bad++;
            }
        if (bad) printf("%s: %d mismatches\n", kNames[op], bad);
        total += bad;
    }
    // every shift multiply 2^0 .. 2^95 (on a quarter of the inputs)
    const u32 np = n / 4;
    int bp = all_pow2(std::make_integer_sequence<int, 96>{}, a, da, dout, np, host);
    if (bp) printf("mul_pow2: %d mismatches\n", bp);
    total += bp;
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dout);
    printf("device bad=%d\n", total);
    return total != 0;
}

// Host check of the Goldilocks primitives in triton-vm-prover_amd/csrc/gl64.hpp
// (the same carry-chain code the gfx950 kernels run) against unsigned __int128
// arithmetic: edge values around p and 2^32, plus random operands.
#include "gl64.hpp"
#include <cstdio>
#include <random>
typedef unsigned __int128 u128;
int main() {
  std::mt19937_64 rng(1);
  const uint64_t P = gl::P; int bad = 0;
  uint64_t edge[] = {0,1,2,P-1,P-2,0xFFFFFFFFull,0x100000000ull,P-0x100000000ull,0xFFFFFFFFFFFFFFFFull,0xFFFFFFFF00000000ull, P+5};
  // 2^-64 mod p (2^64 = 2^32 - 1 mod p)
  const uint64_t RINV = gl::inv(0xFFFFFFFFull);
  for (int it = 0; it < 500000; it++) {
    uint64_t a = it < 121 ? edge[it % 11] : rng(), b = it < 121 ? edge[it / 11] : rng();
    uint64_t ac = a % P, bc = b % P;
    if (gl::add(ac, bc) != (uint64_t)(((u128)ac + bc) % P)) bad++;
    if (gl::sub(ac, bc) != (uint64_t)(((u128)ac + P - bc) % P)) bad++;
    if (gl::canon(a) != a % P) bad++;
    if (gl::mul_lazy(a, b) % P != (uint64_t)(((u128)a * b) % P)) bad++;
    { uint64_t lo, hi; gl::mul_wide_4mad(a, b, lo, hi); u128 w = (u128)a * b;
      if (lo != (uint64_t)w || hi != (uint64_t)(w >> 64)) bad++; }
    { uint64_t lo, hi; gl::sqr_wide(a, lo, hi); u128 w = (u128)a * a;
      if (lo != (uint64_t)w || hi != (uint64_t)(w >> 64)) bad++;
      if (gl::mont_sqr(a) != gl::mont_mul(a, a)) bad++; }
    if (gl::add_lazy(a, bc) % P != (uint64_t)(((u128)a + bc) % P)) bad++;
    {  // Montgomery product of lazy inputs: congruent to a*b*2^-64; canonical if hi < p
      u128 w2 = (u128)a * b;
      uint64_t m = gl::mont_mul(a, b);
      uint64_t want = (uint64_t)((u128)(uint64_t)(w2 % P) * RINV % P);
      if (m % P != want) bad++;
      if ((uint64_t)(w2 >> 64) < P && m >= P) bad++;
      if (gl::montyred(a, 0) != (uint64_t)((u128)(a % P) * RINV % P)) bad++;
      if (gl::to_mont(a) % P != (uint64_t)((u128)a * 0xFFFFFFFFull % P)) bad++;
      if (gl::montyred(gl::to_mont(a), 0) != a % P) bad++;
    }
    u128 w = ((u128)rng() << 64) | rng();
    if (gl::fold128((uint64_t)w, (uint64_t)(w >> 64)) % P != (uint64_t)(w % P)) bad++;
    if (gl::canon(gl::mul_pow2_small<12>(a)) != (uint64_t)(((u128)(a % P) << 12) % P)) bad++;
    if (gl::canon(gl::mul_pow2_small<24>(a)) != (uint64_t)(((u128)(a % P) << 24) % P)) bad++;
    if (gl::canon(gl::mul_pow2_small<32>(a)) != (uint64_t)(((u128)(a % P) << 32) % P)) bad++;
  }
  // mul_pow2<E> for every E < 96 on edge and random operands
  {
    uint64_t xs[64];
    for (int i = 0; i < 11; i++) xs[i] = edge[i];
    for (int i = 11; i < 64; i++) xs[i] = rng();
    auto ref = [&](uint64_t x, int e) { u128 r = x % P; for (int i = 0; i < e; i++) r = (r * 2) % P; return (uint64_t)r; };
    int e = 0;
    auto one = [&](auto f) { for (int i = 0; i < 64; i++) if (f(xs[i]) != ref(xs[i], e)) bad++; e++; };
#define C4(b) one([](uint64_t x){return gl::mul_pow2<b>(x);}); one([](uint64_t x){return gl::mul_pow2<b+1>(x);}); \
              one([](uint64_t x){return gl::mul_pow2<b+2>(x);}); one([](uint64_t x){return gl::mul_pow2<b+3>(x);});
#define C16(b) C4(b) C4(b+4) C4(b+8) C4(b+12)
    C16(0) C16(16) C16(32) C16(48) C16(64) C16(80)
#undef C16
#undef C4
    if (e != 96) bad++;
  }
  // lin52: L + H * 2^32 mod p for L, H < 2^52 (edges and random)
  {
    const uint64_t M52 = (1ull << 52) - 1;
    uint64_t ed[] = {0, 1, 0xFFFFFFFFull, 0x100000000ull, M52, M52 - 1, 0xFFFFF00000000ull, 0xFFFFFFFFFFFFFull,
                     0x100000000000ull, 0xFFFFFFFF0ull};
    for (int i = 0; i < 200000; i++) {
      uint64_t L = i < 100 ? ed[i % 10] : (rng() & M52), H = i < 100 ? ed[i / 10] : (rng() & M52);
      uint64_t v = gl::lin52(L, H);
      if (v % P != (uint64_t)(((u128)L + ((u128)H << 32)) % P)) bad++;
    }
  }
  for (int e = 0; e < 192; e++) { u128 r = 1; for (int i=0;i<e;i++) r = (r*2)%P; if (gl::pow2_mod_p(e) != (uint64_t)r) bad++; }
  printf("bad=%d\n", bad);
  return bad != 0;
}

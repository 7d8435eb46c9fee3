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
  for (int it = 0; it < 500000; it++) {
    uint64_t a = it < 121 ? edge[it % 11] : rng(), b = it < 121 ? edge[it / 11] : rng();
    uint64_t ac = a % P, bc = b % P;
    if (gl::add(ac, bc) != (uint64_t)(((u128)ac + bc) % P)) bad++;
    if (gl::sub(ac, bc) != (uint64_t)(((u128)ac + P - bc) % P)) bad++;
    if (gl::canon(a) != a % P) bad++;
    if (gl::mul_lazy(a, b) % P != (uint64_t)(((u128)a * b) % P)) bad++;
    if (gl::mul_lazy_cc(a, b) % P != (uint64_t)(((u128)a * b) % P)) bad++;
    if (gl::add_lazy(a, bc) % P != (uint64_t)(((u128)a + bc) % P)) bad++;
    u128 w = ((u128)rng() << 64) | rng();
    if (gl::fold128((uint64_t)w, (uint64_t)(w >> 64)) % P != (uint64_t)(w % P)) bad++;
    if (gl::canon(gl::mul_pow2_small<12>(a)) != (uint64_t)(((u128)(a % P) << 12) % P)) bad++;
    if (gl::canon(gl::mul_pow2_small<24>(a)) != (uint64_t)(((u128)(a % P) << 24) % P)) bad++;
    if (gl::canon(gl::mul_pow2_small<32>(a)) != (uint64_t)(((u128)(a % P) << 32) % P)) bad++;
  }
  for (int e = 0; e < 192; e++) { u128 r = 1; for (int i=0;i<e;i++) r = (r*2)%P; if (gl::pow2_mod_p(e) != (uint64_t)r) bad++; }
  printf("bad=%d\n", bad);
  return bad != 0;
}

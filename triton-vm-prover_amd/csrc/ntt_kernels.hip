// Batched Goldilocks NTT and coset low-degree extension for gfx950.
//
// Math restated from src/ntt/ntt.cpp:26-131 and src/lde/lde_randomized.cpp:
// the LDE of a column x of length N onto the coset g*<w_M>, M = E*N, is
//   out[E*q + s] = NTT_N(f_s)[q],  f_s[r] = sum_{i = r mod N} p_i (g w_M^s)^i
// where p = INTT_N(x) * o^-i + (X^N - o^N) * rand(X) (o = trace offset).  The
// E cosets replace the reference's zero-pad + size-M NTT (same exact values,
// a quarter of the butterflies at E = 4).
//
// Each size-N transform is a four-step split N = N1 * N2 (N1 = 2^a):
//   step 1: N1-point DFTs down the columns of the N1 x N2 view, twiddle w_N^(n2 k1)
//   step 2: N2-point DFTs along rows (k1 fixed)
// The LDE fuses INTT step 2, the coset scaling of all E cosets and forward
// step 1 into one kernel (k_lde_step2), so a column is read from HBM three
// times and written three times, with d_out as the only working memory:
//   K1  trace            -> Y   at out[E*N2*k1 + n2]
//   K2  Y (rows k1)      -> Z_s at out[E*(u1 + N2*k1) + s]   (in place per row)
//   K3  Z (u1-tile, all s) -> codewords at out[E*(u1 + N2*u2) + s] (in place)
// Sub-DFTs of length L <= 4096 run in LDS as Stockham radix-16 passes: each
// thread holds R values, applies w_(Ns*R)^(k*r) from a table and an R-point
// DFT whose twiddles are powers of two (w_64 = 2^39), natural order in/out.
#include <atomic>
#include <map>
#include <utility>
#include <mutex>
#include <tuple>

#include "tvh_internal.hpp"

namespace tvh {

constexpr int TPB = 1024;    // threads per block of every NTT kernel (16 waves)
constexpr int TILE = 16384;  // u64 elements per block: 16 per thread, 128 KB before padding
constexpr int LOG_TILE = 14;

__device__ __forceinline__ int pidx(int t, int i, int Lp) { return t * Lp + i + (i >> 4); }
__host__ __device__ inline int padded_len(int L) { return L + (L >> 4) + 1; }

template <int LOGR>
__device__ __forceinline__ constexpr int brev(int i) {
    int r = 0;
    for (int b = 0; b < LOGR; b++)
        if (i & (1 << b)) r |= 1 << (LOGR - 1 - b);
    return r;
}

// R-point DFT in registers, natural order in and out: radix-2 DIT on the
// bit-reversed input.  The butterflies of stage LEN use w_LEN^j =
// 2^(39*(64/LEN)*j mod 192) (w_64 = 2^39); a twiddle 2^e with e >= 96 is
// -2^(e-96), folded into the butterfly by swapping add and sub; 2^e itself is
// applied by word shifts (gl::mul_pow2).  Every index and exponent is a
// template constant (index sequences), so the whole DFT is straight-line code.
template <bool INV, int LEN, int J>
__device__ __forceinline__ void dft_bfly(u64& x, u64& y) {
    constexpr int e0 = (39 * (64 / LEN) * J) % 192;
    constexpr int e = INV ? (192 - e0) % 192 : e0;
    constexpr bool negw = e >= 96;
    constexpr int ep = negw ? e - 96 : e;
    u64 t = ep == 0 ? y : gl::mul_pow2<ep>(y);
    u64 u = x;
    if (negw) {
        x = gl::sub(u, t);
        y = gl::add(u, t);
    } else {
        x = gl::add(u, t);
        y = gl::sub(u, t);
    }
}

template <bool INV, int LEN, size_t... K>
__device__ __forceinline__ void dft_stage(u64* a, std::index_sequence<K...>) {
    // butterfly K: group K / (LEN/2), offset j = K % (LEN/2)
    (dft_bfly<INV, LEN, (int)(K % (LEN / 2))>(a[(K / (LEN / 2)) * LEN + K % (LEN / 2)],
                                             a[(K / (LEN / 2)) * LEN + K % (LEN / 2) + LEN / 2]),
     ...);
}

template <int LOGR, bool INV, size_t... S>
__device__ __forceinline__ void dft_stages(u64* a, std::index_sequence<S...>) {
    (dft_stage<INV, (2 << S)>(a, std::make_index_sequence<(1 << LOGR) / 2>{}), ...);
}

template <int LOGR, bool INV>
__device__ __forceinline__ void dft_reg(u64* v) {
    constexpr int R = 1 << LOGR;
    u64 a[R];
#pragma unroll
    for (int i = 0; i < R; i++) a[i] = v[brev<LOGR>(i)];
    dft_stages<LOGR, INV>(a, std::make_index_sequence<LOGR>{});
#pragma unroll
    for (int i = 0; i < R; i++) v[i] = a[i];
}

template <int LOGR, bool INV, int NTPB = TPB, int NTILE = TILE>
__device__ __forceinline__ void stockham_pass(u64* buf, int logL, int Lp, int NT, int Ns,
                                              const u64* __restrict__ r4096) {
    constexpr int R = 1 << LOGR;
    constexpr int MAXIT = NTILE / (R * NTPB);
    const int per = 1 << (logL - LOGR);
    const int total = NT << (logL - LOGR);
    const int step = 4096 / (Ns * R);
    static_assert(MAXIT >= 1, "radix larger than the per-thread element count");
    u64 v[MAXIT][R];
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        int w = threadIdx.x + it * NTPB;
        if (w < total) {
            int t = w >> (logL - LOGR);
            int j = w & (per - 1);
            int k = j & (Ns - 1);
#pragma unroll
            for (int r = 0; r < R; r++) v[it][r] = buf[pidx(t, j + r * per, Lp)];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[it][r] = gl::mont_mul(v[it][r], r4096[k * r * step]);
            }
            dft_reg<LOGR, INV>(v[it]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        int w = threadIdx.x + it * NTPB;
        if (w < total) {
            int j = (w) & (per - 1);
            int t = w >> (logL - LOGR);
            int k = j & (Ns - 1);
            int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) buf[pidx(t, base + r * Ns, Lp)] = v[it][r];
        }
    }
    __syncthreads();
}

// NT transforms of length 2^logL (<= 4096) in LDS; NT * L <= NTILE; NTPB
// threads hold NTILE / NTPB values each.
template <bool INV, int NTPB = TPB, int NTILE = TILE>
__device__ __forceinline__ void lds_dft(u64* buf, int logL, int NT, const u64* __restrict__ r4096) {
    const int Lp = padded_len(1 << logL);
    int Ns = 1;
    int rem = logL;
    while (rem >= 4) {
        stockham_pass<4, INV, NTPB, NTILE>(buf, logL, Lp, NT, Ns, r4096);
        Ns <<= 4;
        rem -= 4;
    }
    if (rem == 3) stockham_pass<3, INV, NTPB, NTILE>(buf, logL, Lp, NT, Ns, r4096);
    else if (rem == 2) stockham_pass<2, INV, NTPB, NTILE>(buf, logL, Lp, NT, Ns, r4096);
    else if (rem == 1) stockham_pass<1, INV, NTPB, NTILE>(buf, logL, Lp, NT, Ns, r4096);
}

// ---------------------------------------------------------------- fused-endpoint DFT
// The same Stockham passes as lds_dft, but the first pass takes its inputs
// straight from global memory (a Load functor) and the last pass writes its
// outputs straight to global memory (a Store functor): per transform one LDS
// write and read per middle pass instead of a global->LDS copy before and an
// LDS->global copy after, and 2 barriers per middle pass only.  A one-pass
// transform (L <= 16) never touches LDS.
//
// Work item (t, j) of a radix-R pass: v[r] = x_t[j + r*per] (times the
// Stockham twiddle when Ns > 1), R-point DFT, y_t[(j - k)*R + k + r*Ns] = v[r]
// with k = j mod Ns.  Functors: load(t, j, per, v) fills v[0..R); store(t, j,
// Ns, v) writes it.  TF (transform-fastest): adjacent threads take adjacent
// transforms (column kernels, whose transforms are adjacent in memory);
// otherwise adjacent elements of one transform (row kernels).
struct LdsIO {
    u64* buf;
    int Lp;
    template <int R>
    __device__ __forceinline__ void load(int t, int j, int per, u64 (&v)[R]) const {
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = buf[pidx(t, j + r * per, Lp)];
    }
    template <int R>
    __device__ __forceinline__ void store(int t, int j, int Ns, const u64 (&v)[R]) const {
        int k = j & (Ns - 1);
        int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) buf[pidx(t, base + r * Ns, Lp)] = v[r];
    }
};

// Workgroup barrier for LDS hand-offs only: waits for this wave's LDS
// operations, not for its global loads (the __syncthreads fence also drains
// the vector-memory counter, which would expose the latency of twiddle loads
// issued ahead of the barrier).  No thread of these kernels reads global
// memory that another thread of the block writes before the barrier.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Stockham twiddles of one pass for the thread's work items, loaded ahead of
// the pass: tw.t[it][r] = r4096[k * r * step] (r >= 1) for work item it.
template <int LOGR, int NTPB, int NTILE>
struct PassTw {
    static constexpr int R = 1 << LOGR;
    static constexpr int MAXIT = NTILE / (R * NTPB);
    u64 t[MAXIT][R];
};

template <int LOGR, int NTPB, int NTILE, bool TF>
__device__ __forceinline__ void prefetch_tw(PassTw<LOGR, NTPB, NTILE>& tw, int logL, int NT,
                                            int logNT, int Ns, const u64* __restrict__ r4096) {
    constexpr int R = 1 << LOGR;
    constexpr int MAXIT = NTILE / (R * NTPB);
    const int lp = logL - LOGR;
    const int per = 1 << lp;
    const int total = NT << lp;
    const int step = 4096 / (Ns * R);
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        int w = threadIdx.x + it * NTPB;
        if (w < total) {
            int j = TF ? (w >> logNT) : (w & (per - 1));
            int k = j & (Ns - 1);
#pragma unroll
            for (int r = 1; r < R; r++) tw.t[it][r] = r4096[k * r * step];
        }
    }
}

template <int LOGR, bool INV, int NTPB, int NTILE, bool TF, bool FROM_LDS, bool TO_LDS, class Ld,
          class St>
__device__ __forceinline__ void fused_pass(const LdsIO& io, int logL, int NT, int logNT, int Ns,
                                           const PassTw<LOGR, NTPB, NTILE>& tw, const Ld& ld,
                                           const St& st) {
    constexpr int R = 1 << LOGR;
    constexpr int MAXIT = NTILE / (R * NTPB);
    static_assert(MAXIT >= 1, "radix larger than the per-thread element count");
    const int lp = logL - LOGR;
    const int per = 1 << lp;
    const int total = NT << lp;
    u64 v[MAXIT][R];
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        int w = threadIdx.x + it * NTPB;
        if (w < total) {
            int t = TF ? (w & (NT - 1)) : (w >> lp);
            int j = TF ? (w >> logNT) : (w & (per - 1));
            if (FROM_LDS) io.load<R>(t, j, per, v[it]);
            else ld.template load<R>(t, j, per, v[it]);
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[it][r] = gl::mont_mul(v[it][r], tw.t[it][r]);
            }
            dft_reg<LOGR, INV>(v[it]);
        }
    }
    if (FROM_LDS && TO_LDS) lds_barrier();  // in place: every read of the pass before any write
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        int w = threadIdx.x + it * NTPB;
        if (w < total) {
            int t = TF ? (w & (NT - 1)) : (w >> lp);
            int j = TF ? (w >> logNT) : (w & (per - 1));
            if (TO_LDS) io.store<R>(t, j, Ns, v[it]);
            else st.template store<R>(t, j, Ns, v[it]);
        }
    }
}

// Straight-line pass sequence for L = 2^(4Q + REM): Q radix-16 passes, then a
// radix-2^REM pass (a radix-1 copy when L = 1).  Every radix, stride and
// first/last role is a compile-time constant, which keeps register pressure
// at one pass's worth.  The twiddles of pass P + 1 are loaded before the
// barrier that ends pass P, so their latency overlaps the barrier wait.
template <int P, int NP, int Q, int REM, bool INV, int NTPB, int NTILE, bool TF, class Ld,
          class St>
__device__ __forceinline__ void fused_chain(const LdsIO& io, int NT, int logNT,
                                            const u64* __restrict__ r4096, const Ld& ld,
                                            const St& st,
                                            const PassTw<(P < Q ? 4 : REM), NTPB, NTILE>& tw) {
    constexpr int logL = 4 * Q + REM;
    fused_pass<(P < Q ? 4 : REM), INV, NTPB, NTILE, TF, (P > 0), (P + 1 < NP)>(
        io, logL, NT, logNT, 1 << (4 * P), tw, ld, st);
    if constexpr (P + 1 < NP) {
        PassTw<(P + 1 < Q ? 4 : REM), NTPB, NTILE> next;
        prefetch_tw<(P + 1 < Q ? 4 : REM), NTPB, NTILE, TF>(next, logL, NT, logNT, 1 << (4 * (P + 1)),
                                                           r4096);
        lds_barrier();
        fused_chain<P + 1, NP, Q, REM, INV, NTPB, NTILE, TF>(io, NT, logNT, r4096, ld, st, next);
    }
}

template <int Q, int REM, bool INV, int NTPB, int NTILE, bool TF, class Ld, class St>
__device__ __forceinline__ void fused_passes(const LdsIO& io, int NT, int logNT,
                                             const u64* __restrict__ r4096, const Ld& ld,
                                             const St& st) {
    constexpr int NP = Q + ((REM != 0 || Q == 0) ? 1 : 0);
    PassTw<(0 < Q ? 4 : REM), NTPB, NTILE> none;  // the first pass has no twiddles
    fused_chain<0, NP, Q, REM, INV, NTPB, NTILE, TF>(io, NT, logNT, r4096, ld, st, none);
}

// LDS row pitch of a transform for the fused DFT: L + L/16 + pad, with pad
// chosen so that the lanes of a 16-lane LDS cycle hit distinct banks under the
// transform-fastest mapping (lane = j * NT' + t, NT' = min(NT, 16)):
// t * pitch + j is then distinct mod 16 for t < NT', j < 16 / NT'.
__host__ __device__ inline int fused_pitch(int L, int NT, bool tf) {
    int ntp = NT < 16 ? NT : 16;
    return L + (L >> 4) + (tf ? 16 / ntp : 1);
}

// NT (power of two) transforms of length 2^logL (<= 4096), NT * L <= NTILE.
// The block's LDS (buf, NT * fused_pitch(L, NT, TF) u64) must not be in use by
// other threads on entry.
template <bool INV, int NTPB, int NTILE, bool TF, class Ld, class St>
__device__ __forceinline__ void fused_dft(u64* buf, int logL, int NT, int logNT,
                                          const u64* __restrict__ r4096, const Ld& ld,
                                          const St& st) {
    const LdsIO io{buf, fused_pitch(1 << logL, NT, TF)};
#define TVH_FUSED(q, r) fused_passes<q, r, INV, NTPB, NTILE, TF>(io, NT, logNT, r4096, ld, st)
    switch (logL) {
        case 0: TVH_FUSED(0, 0); break;
        case 1: TVH_FUSED(0, 1); break;
        case 2: TVH_FUSED(0, 2); break;
        case 3: TVH_FUSED(0, 3); break;
        case 4: TVH_FUSED(1, 0); break;
        case 5: TVH_FUSED(1, 1); break;
        case 6: TVH_FUSED(1, 2); break;
        case 7: TVH_FUSED(1, 3); break;
        case 8: TVH_FUSED(2, 0); break;
        case 9: TVH_FUSED(2, 1); break;
        case 10: TVH_FUSED(2, 2); break;
        case 11: TVH_FUSED(2, 3); break;
        default: TVH_FUSED(3, 0); break;  // 12
    }
#undef TVH_FUSED
}

// w_N^(+-x) for x < N via the two-level 2^24 tables, in Montgomery form (the
// tables hold w * 2^64 mod p): apply with gl::mont_mul.
__device__ __forceinline__ u64 twiddle(const Roots& rt, int dir, u32 x, u32 logN) {
    u32 y = x << (kMaxLogN - logN);
    return gl::mont_mul(rt.lo[dir][y & 4095], rt.hi[dir][y >> 12]);
}

// Per-block segment lookup for multi-table launches: start[i] is the first
// block of segment i (start[nseg] = total blocks); returns the segment of blk.
__device__ __forceinline__ u32 find_segment(const u32* __restrict__ start, u32 nseg, u32 blk) {
    u32 lo = 0, hi = nseg;  // invariant: start[lo] <= blk < start[hi]
    while (hi - lo > 1) {
        u32 mid = (lo + hi) >> 1;
        if (start[mid] <= blk) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Transforms s = (c, n2): element i at src[c*sstride + (i << b) + n2].
// Addressing in the column and row functors: every access is a block-uniform
// base (the first transform of the block) plus a 32-bit element offset, so the
// loads and stores take a scalar base and a vector offset and no 64-bit
// address arithmetic runs per element.  Offsets stay below 2^32 elements: a
// block's transforms span at most its tile's worth of columns, and the
// in-column offset is below M <= 2^27.
struct Step1Load {
    const u64* src;
    u64 sstride, s0, ntr;
    u32 b;
    template <int R>
    __device__ __forceinline__ void load(int t, int j, int per, u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s < ntr) {
            const u64 mask = (1ull << b) - 1;
            const u64* base = src + (s0 >> b) * sstride + (s0 & mask);
            const u32 o = (u32)(((s >> b) - (s0 >> b)) * sstride + (s & mask) - (s0 & mask));
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = base[o + ((u32)(j + r * per) << b)];
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = 0;
        }
    }
};
// Output k1 of transform (c, n2), times w_N^(+-n2*k1), to dst[c*dstride + E*N2*k1 + n2].
template <bool INV>
struct Step1Store {
    u64* dst;
    u64 dstride, s0, ntr;
    u32 b, E, logN;
    Roots rt;
    template <int R>
    __device__ __forceinline__ void store(int t, int j, int Ns, const u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s >= ntr) return;
        const u64 mask = (1ull << b) - 1;
        u64 n2 = s & mask;
        u64* p0 = dst + (s0 >> b) * dstride + (s0 & mask);
        const u32 o = (u32)(((s >> b) - (s0 >> b)) * dstride + n2 - (s0 & mask));
        int k = j & (Ns - 1), base = (j - k) * R + k;
        // table entries first (index 0 holds 1, so n2 = 0 or k1 = 0 needs no
        // branch), then the products: the loads are in flight together
        const u32 sh = kMaxLogN - logN;
        u64 tl[R], th[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            u32 y = (u32)(n2 * (u32)(base + r * Ns)) << sh;
            tl[r] = rt.lo[INV][y & 4095];
            th[r] = rt.hi[INV][y >> 12];
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            u32 k1 = (u32)(base + r * Ns);
            p0[o + ((E * k1) << b)] = gl::mont_mul(v[r], gl::mont_mul(tl[r], th[r]));
        }
    }
};

// ---------------------------------------------------------------- step 1
// Transforms s = (c, n2), element n1 at src[c*sstride + N2*n1 + n2]; DFT over n1
// (inverse if INV) -> k1; times w_N^(+-n2*k1); to dst[c*dstride + E*N2*k1 + n2].
// Adjacent threads take adjacent transforms (adjacent n2: contiguous in HBM).
template <bool INV, int LT = LOG_TILE>
__device__ __forceinline__ void step1_body(const Step1Args& A, u64 blk, const Roots& rt) {
    extern __shared__ u64 lds[];
    constexpr int NTPB = 1 << (LT - 4);
    const u32 b = A.logN - A.a;
    const u64 s0 = blk << A.logNT;
    const Step1Load ld{A.src, A.sstride, s0, A.ntr, b};
    const Step1Store<INV> st{A.dst, A.dstride, s0, A.ntr, b, A.E, A.logN, rt};
    fused_dft<INV, NTPB, 1 << LT, true>(lds, A.a, 1 << A.logNT, A.logNT, rt.r4096[INV], ld, st);
}

template <bool INV, int LT = LOG_TILE>
__global__ void __launch_bounds__(1 << (LT - 4), 4) k_step1(Step1Args A, Roots rt) {
    step1_body<INV, LT>(A, blockIdx.x, rt);
}

template <bool INV, int LT = LOG_TILE>
__global__ void __launch_bounds__(1 << (LT - 4), 4) k_step1_multi(const Step1Args* __restrict__ args,
                                                                  const u32* __restrict__ start,
                                                                  u32 nseg, Roots rt) {
    u32 seg = find_segment(start, nseg, blockIdx.x);
    const Step1Args& A = args[seg];
    step1_body<INV, LT>(A, blockIdx.x - start[seg], rt);
}

// ---------------------------------------------------------------- plain NTT step 2
// Rows s = (c, k1): n2 at src[c*N + N2*k1 + n2]; DFT over n2 -> k2 (times N^-1
// when INV); to dst[c*N + k1 + N1*k2].
template <bool INV>
__global__ void __launch_bounds__(TPB) k_ntt_step2(const u64* __restrict__ src,
                                                   u64* __restrict__ dst, u32 logN, u32 a,
                                                   u64 nrows, u32 logNT, u64 ninv, Roots rt) {
    extern __shared__ u64 lds[];
    const u32 b = logN - a;
    const int L = 1 << b;
    const int Lp = padded_len(L);
    const int NT = 1 << logNT;
    const u64 s0 = (u64)blockIdx.x << logNT;
    for (int e = threadIdx.x; e < NT * L; e += TPB) {
        int t = e >> b, i = e & (L - 1);
        u64 s = s0 + t;
        if (s < nrows) lds[pidx(t, i, Lp)] = src[(s << b) + i];
    }
    __syncthreads();
    lds_dft<INV>(lds, b, NT, rt.r4096[INV]);
    for (int e = threadIdx.x; e < NT * L; e += TPB) {
        int t = e & (NT - 1), k2 = e >> logNT;
        u64 s = s0 + t;
        if (s < nrows) {
            u64 c = s >> a, k1 = s & ((1ull << a) - 1);
            u64 v = lds[pidx(t, k2, Lp)];
            if (INV) v = gl::mul(v, ninv);
            dst[(c << logN) + k1 + ((u64)k2 << a)] = v;
        }
    }
}

// ---------------------------------------------------------------- LDE step 2
__device__ __forceinline__ u64 rand_term(const LdeArgs& A, u64 c, u32 s, u64 r) {
    const RandSpec& rs = A.rs;
    u32 limb = (u32)(c % rs.group);
    if (!((rs.limb_mask >> limb) & 1)) return 0;
    const u64* p = rs.ptr + (c / rs.group) * rs.group_stride + limb * rs.limb_stride +
                   r * rs.elem_stride;
    // select delta_n[s] without a dynamic index (a dynamic index into the
    // by-value kernel argument would copy it to scratch)
    u64 d = A.delta_n[0];
#pragma unroll
    for (u32 i = 1; i < 16; i++)
        if (s == i) d = A.delta_n[i];
    return gl::mul(gl::mul(*p, d), gl::pow(A.toff, r));
}

// INTT = false: the rows already hold the inverse-transformed coefficients
// (written by k_lde_rowinv), only the coset expansion runs here.
template <bool INTT = true>
__device__ __forceinline__ void lde_step2_body(const LdeArgs& A, u64 blk, const Roots& rt) {
    extern __shared__ u64 lds[];
    u64* out = A.out;
    const u32 a = A.a, b = A.logN - A.a, E = A.E, logNT = A.logNT;
    const int L = 1 << b;
    const int Lp = padded_len(L);
    const int NT = 1 << logNT;           // rows per block
    const u64 nrows = A.cols << a;
    const u64 s0 = blk << logNT;
    // load Y rows into transforms 0..NT-1 (slot s = 0)
    for (int e = threadIdx.x; e < NT * L; e += TPB) {
        int t = e >> b, i = e & (L - 1);
        u64 s = s0 + t;
        if (s < nrows) {
            u64 c = s >> a, k1 = s & ((1ull << a) - 1);
            lds[pidx(t, i, Lp)] = A.src[c * A.src_col_stride + k1 * A.src_row_stride + i];
        }
    }
    __syncthreads();
    if (INTT) lds_dft<true>(lds, b, NT, rt.r4096[1]);
    // expand: transform (sl*NT + t) <- f_sl for row t
    for (int e = threadIdx.x; e < NT * L; e += TPB) {
        int t = e >> b, k2 = e & (L - 1);
        u64 s = s0 + t;
        if (s < nrows) {
            u64 c = s >> a, k1 = s & ((1ull << a) - 1);
            u64 C = lds[pidx(t, k2, Lp)];
            u64 r = k1 + ((u64)k2 << a);
            for (u32 sl = 0; sl < E; sl++) {
                u64 v = C;
                if (r < A.rs.k) v = gl::add(v, rand_term(A, c, sl, r));
                lds[pidx(sl * NT + t, k2, Lp)] = gl::mont_mul(v, A.pb[(sl << b) + k2]);
            }
        }
    }
    __syncthreads();
    lds_dft<false>(lds, b, NT * E, rt.r4096[0]);
    // write Z_s[u1][k1] to out[E*(u1 + N2*k1) + s], contiguous over (u1, s)
    // Output factor pa[s][k1] * w_N^(k1*u1) (N^-1 beta_s^k1 and the four-step
    // twiddle; pb was applied before the row DFT).
    const int logRW = b + A.logE;  // row of all cosets: E * N2 elements
    const int RW = 1 << logRW;
    if (RW == TILE) {
        // one row per block (k1 uniform): thread outputs q = tid + TPB*j have
        // the fixed coset sl = tid mod E and u1 = tid/E + (TPB/E)*j, so the
        // factor advances by w_N^(k1*TPB/E) per j
        u64 s = s0;
        if (s < nrows) {
            u64 c = s >> a, k1 = s & ((1ull << a) - 1);
            const int sl = threadIdx.x & (E - 1);
            const u32 u0 = threadIdx.x >> A.logE;
            u64 f = A.pa[(sl << a) + k1];
            if (k1 && u0) f = gl::mont_mul(f, twiddle(rt, 0, (u32)(k1 * u0), A.logN));
            const u64 step = twiddle(rt, 0, (u32)(k1 * (TPB >> A.logE)), A.logN);
            u64* dst = out + c * A.M + (((u64)E * k1) << b);
#pragma unroll
            for (int j = 0; j < TILE / TPB; j++) {
                int q = threadIdx.x + j * TPB;
                int u1 = q >> A.logE;
                dst[q] = gl::mont_mul(lds[pidx(sl, u1, Lp)], f);
                if (j + 1 < TILE / TPB) f = gl::mont_mul(f, step);
            }
        }
    } else {
        for (int e = threadIdx.x; e < NT * RW; e += TPB) {
            int t = e >> logRW, q = e & (RW - 1);
            int u1 = q >> A.logE, sl = q & (E - 1);
            u64 s = s0 + t;
            if (s < nrows) {
                u64 c = s >> a, k1 = s & ((1ull << a) - 1);
                u64 f = A.pa[(sl << a) + k1];
                if (u1 && k1) f = gl::mont_mul(f, twiddle(rt, 0, (u32)(k1 * u1), A.logN));
                out[c * A.M + (((u64)E * k1) << b) + q] =
                    gl::mont_mul(lds[pidx(sl * NT + t, u1, Lp)], f);
            }
        }
    }
}

__global__ void __launch_bounds__(TPB) k_lde_step2(LdeArgs A, Roots rt) {
    lde_step2_body<true>(A, blockIdx.x, rt);
}

// Split form of the fused row step: k_lde_rowinv inverse-transforms TILE / N2
// rows per block with every thread busy and writes the coefficients where the
// row was (out[c*M + E*N2*k1 + n2]); k_lde_expand then runs the coset
// expansion of k_lde_step2 on them.  One extra read and write of N elements
// per column buys full occupancy for the inverse row DFT, which the fused
// kernel runs with a quarter of its threads at E = 4.
// Rows s = (c, k1) of N2 elements at src[c*scs + k1*srs + i].
struct RowLoad {
    const u64* src;
    u64 scs, srs, s0, nrows;
    u32 a;
    template <int R>
    __device__ __forceinline__ void load(int t, int j, int per, u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s < nrows) {
            const u64 mask = (1ull << a) - 1;
            const u64* p = src + (s0 >> a) * scs + (s0 & mask) * srs;
            const u32 o = (u32)(((s >> a) - (s0 >> a)) * scs + ((s & mask) - (s0 & mask)) * srs);
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = p[o + (u32)(j + r * per)];
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = 0;
        }
    }
};
// Row (c, k1) to out[c*M + E*N2*k1 + i].
struct RowStore {
    u64* out;
    u64 M, s0, nrows;
    u32 a, b, E;
    template <int R>
    __device__ __forceinline__ void store(int t, int j, int Ns, const u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s >= nrows) return;
        const u64 mask = (1ull << a) - 1;
        u64* p = out + (s0 >> a) * M + (((u64)E * (s0 & mask)) << b);
        const u32 o = (u32)(((s >> a) - (s0 >> a)) * M + (((u64)E * ((s & mask) - (s0 & mask))) << b));
        int k = j & (Ns - 1), base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) p[o + (u32)(base + r * Ns)] = v[r];
    }
};

// Row tiles of 2^LT elements, 16 per thread (LT = 14: 1024 threads, one block
// per CU; LT = 13: 512 threads, two blocks per CU).
template <int LT = LOG_TILE>
__device__ __forceinline__ void lde_rowinv_body(const LdeArgs& A, u64 blk, const Roots& rt) {
    extern __shared__ u64 lds[];
    const u32 b = A.logN - A.a;
    const u64 nrows = A.cols << A.a;
    const u64 s0 = blk << A.logNT;  // rows per block of this kernel: 2^LT / N2
    const RowLoad ld{A.src, A.src_col_stride, A.src_row_stride, s0, nrows, A.a};
    const RowStore st{A.out, A.M, s0, nrows, A.a, b, A.E};
    fused_dft<true, (1 << (LT - 4)), (1 << LT), false>(lds, b, 1 << A.logNT, A.logNT,
                                                        rt.r4096[1], ld, st);
}

template <int LT>
__global__ void __launch_bounds__(1 << (LT - 4)) k_lde_rowinv(LdeArgs A, Roots rt) {
    lde_rowinv_body<LT>(A, blockIdx.x, rt);
}

template <int LT>
__global__ void __launch_bounds__(1 << (LT - 4)) k_lde_rowinv_multi(const LdeArgs* __restrict__ args,
                                                                    const u32* __restrict__ start,
                                                                    u32 nseg, Roots rt) {
    u32 seg = find_segment(start, nseg, blockIdx.x);
    const LdeArgs& A = args[seg];
    lde_rowinv_body<LT>(A, blockIdx.x - start[seg], rt);
}

// Coset transforms T = t*E + sl of row t (coefficients at src rows, written
// by k_lde_rowinv): element k2 = (C_k2 + randomizer term) * beta_sl^(N1*k2).
// Cosets are the fastest index, so that with the transform-fastest thread
// mapping the E cosets of one output position are written by adjacent lanes
// (out[E*u1 + sl] contiguous).
struct ExpandLoad {
    const LdeArgs& A;
    u64 s0, nrows;
    template <int R>
    __device__ __forceinline__ void load(int T, int j, int per, u64 (&v)[R]) const {
        const u32 a = A.a, b = A.logN - A.a;
        u32 sl = (u32)T & (A.E - 1);
        u64 s = s0 + ((u32)T >> A.logE);
        if (s >= nrows) {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = 0;
            return;
        }
        const u64 mask = (1ull << a) - 1;
        u64 c = s >> a, k1 = s & mask;
        const u64* p = A.src + (s0 >> a) * A.src_col_stride + (s0 & mask) * A.src_row_stride;
        const u32 o = (u32)((c - (s0 >> a)) * A.src_col_stride + (k1 - (s0 & mask)) * A.src_row_stride);
        const u32 ob = sl << b;
        // every load first, with no branch in between, so that they are in
        // flight together
        u64 f[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            v[r] = p[o + (u32)(j + r * per)];
            f[r] = A.pb[ob + (u32)(j + r * per)];
        }
        // randomizer terms: element k2 of row k1 is trace row k1 + k2 * N1,
        // which carries one when below rand_len (possible only for k1 < rand_len)
        if (k1 < A.rs.k) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                u64 rr = k1 + ((u64)(j + r * per) << a);
                if (rr < A.rs.k) v[r] = gl::add(v[r], rand_term(A, c, sl, rr));
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = gl::mont_mul(v[r], f[r]);
    }
};
// Output u1 of coset sl, row k1, times N^-1 beta_sl^k1 w_N^(k1*u1), to
// out[c*M + E*N2*k1 + E*u1 + sl].  On the last radix-16 pass of a full row the
// thread's outputs are u1 = j + Ns*r, so the factor advances by w_N^(k1*Ns).
struct ExpandStore {
    const LdeArgs& A;
    Roots rt;
    u64 s0, nrows;
    template <int R>
    __device__ __forceinline__ void store(int T, int j, int Ns, const u64 (&v)[R]) const {
        const u32 a = A.a, b = A.logN - A.a, E = A.E;
        u32 sl = (u32)T & (E - 1);
        u64 s = s0 + ((u32)T >> A.logE);
        if (s >= nrows) return;
        u64 k1 = s & ((1ull << a) - 1);
        // row s starts at out + s*E*N2 (c*M + k1*E*N2 with M = N1*E*N2)
        u64* dst = A.out + ((s0 * E) << b);
        const u32 o = (u32)(((s - s0) * E) << b) + sl;
        // (twiddle index 0 holds 1: no branch for k1 = 0 or u1 = 0)
        u64 f = A.pa[((u64)sl << a) + k1];
        int k = j & (Ns - 1), base = (j - k) * R + k;
        if (R * Ns == (1 << b)) {  // base = j: u1 = j + Ns*r
            f = gl::mont_mul(f, twiddle(rt, 0, (u32)(k1 * j), A.logN));
            const u64 step = twiddle(rt, 0, (u32)(k1 * Ns), A.logN);
#pragma unroll
            for (int r = 0; r < R; r++) {
                dst[o + E * (u32)(j + r * Ns)] = gl::mont_mul(v[r], f);
                if (r + 1 < R) f = gl::mont_mul(f, step);
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) {
                u32 u1 = (u32)(base + r * Ns);
                u64 g = gl::mont_mul(f, twiddle(rt, 0, (u32)(k1 * u1), A.logN));
                dst[o + E * u1] = gl::mont_mul(v[r], g);
            }
        }
    }
};

// Split form, second kernel: for NT rows per block (NT * E * N2 = 2^LT), the E
// coset transforms straight from the coefficient rows to the codewords.  The
// outputs overwrite the coefficient row in place, so every coset of a row
// must be computed by the block that reads the row (its reads all precede its
// first barrier, its writes follow its last one).
template <int LT = LOG_TILE>
__device__ __forceinline__ void lde_expand_body(const LdeArgs& A, u64 blk, const Roots& rt) {
    extern __shared__ u64 lds[];
    const u32 b = A.logN - A.a;
    const u64 nrows = A.cols << A.a;
    const u64 s0 = blk << A.logNT;
    const ExpandLoad ld{A, s0, nrows};
    const ExpandStore st{A, rt, s0, nrows};
    fused_dft<false, (1 << (LT - 4)), (1 << LT), true>(
        lds, b, (int)(A.E << A.logNT), (int)(A.logE + A.logNT), rt.r4096[0], ld, st);
}

template <int LT>
__global__ void __launch_bounds__(1 << (LT - 4)) k_lde_expand(LdeArgs A, Roots rt) {
    lde_expand_body<LT>(A, blockIdx.x, rt);
}

template <int LT>
__global__ void __launch_bounds__(1 << (LT - 4)) k_lde_expand_multi(const LdeArgs* __restrict__ args,
                                                                    const u32* __restrict__ start,
                                                                    u32 nseg, Roots rt) {
    u32 seg = find_segment(start, nseg, blockIdx.x);
    const LdeArgs& A = args[seg];
    lde_expand_body<LT>(A, blockIdx.x - start[seg], rt);
}

__global__ void __launch_bounds__(TPB) k_lde_step2_multi(const LdeArgs* __restrict__ args,
                                                         const u32* __restrict__ start, u32 nseg,
                                                         Roots rt) {
    u32 seg = find_segment(start, nseg, blockIdx.x);
    const LdeArgs& A = args[seg];  // read in place: a private copy would spill delta_n
    lde_step2_body<true>(A, blockIdx.x - start[seg], rt);
}

// ---------------------------------------------------------------- LDE step 3
// Transforms (c, q = u1*E + s), element k1 at out[c*M + E*N2*k1 + q]; forward
// DFT over k1 -> u2, written back to out[c*M + E*N2*u2 + q].
// Transforms (c, q) of one column: element i at out[c*M + (i << logQ) + q], in place.
struct Step3IO {
    u64* out;
    u64 M, s0, ntr;
    u32 logQ;
    __device__ __forceinline__ u64* base() const {
        return out + (s0 >> logQ) * M + (s0 & ((1ull << logQ) - 1));
    }
    __device__ __forceinline__ u32 off(u64 s) const {
        const u64 mask = (1ull << logQ) - 1;
        return (u32)(((s >> logQ) - (s0 >> logQ)) * M + (s & mask) - (s0 & mask));
    }
    template <int R>
    __device__ __forceinline__ void load(int t, int j, int per, u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s < ntr) {
            const u64* p = base();
            const u32 o = off(s);
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = p[o + ((u32)(j + r * per) << logQ)];
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = 0;
        }
    }
    template <int R>
    __device__ __forceinline__ void store(int t, int j, int Ns, const u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s >= ntr) return;
        u64* p = base();
        const u32 o = off(s);
        int k = j & (Ns - 1), base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) p[o + ((u32)(base + r * Ns) << logQ)] = v[r];
    }
};

template <int LT = LOG_TILE>
__device__ __forceinline__ void lde_step3_body(const Step3Args& A, u64 blk, const Roots& rt) {
    extern __shared__ u64 lds[];
    constexpr int NTPB = 1 << (LT - 4);
    const u32 b = A.logN - A.a;
    const u32 logQ = b + A.logE;  // transforms per column = E*N2
    const Step3IO io{A.out, A.M, blk << A.logNT, A.cols << logQ, logQ};
    // every input of the block is read in the first pass, before any output is
    // written: in place is safe
    fused_dft<false, NTPB, 1 << LT, true>(lds, A.a, 1 << A.logNT, A.logNT, rt.r4096[0], io, io);
}

template <int LT = LOG_TILE>
__global__ void __launch_bounds__(1 << (LT - 4), 4) k_lde_step3(Step3Args A, Roots rt) {
    lde_step3_body<LT>(A, blockIdx.x, rt);
}

template <int LT = LOG_TILE>
__global__ void __launch_bounds__(1 << (LT - 4), 4) k_lde_step3_multi(const Step3Args* __restrict__ args,
                                                                      const u32* __restrict__ start,
                                                                      u32 nseg, Roots rt) {
    u32 seg = find_segment(start, nseg, blockIdx.x);
    const Step3Args& A = args[seg];
    lde_step3_body<LT>(A, blockIdx.x - start[seg], rt);
}

// Constant tables pa/pb of one (logN, a, E, o, g).
__global__ void k_lde_consts(u64* pa, u64* pb, u32 a, u32 b, u32 E, const u64* beta, u64 ninv) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    u64 n1 = 1ull << a, n2 = 1ull << b;
    // stored in Montgomery form (x * 2^64 mod p) for gl::mont_mul
    if (i < E * n1) {
        u64 s = i >> a, k1 = i & (n1 - 1);
        pa[i] = gl::canon(gl::to_mont(gl::mul(ninv, gl::pow(beta[s], k1))));
    } else if (i < E * (n1 + n2)) {
        u64 j = i - E * n1;
        u64 s = j >> b, k2 = j & (n2 - 1);
        pb[j] = gl::canon(gl::to_mont(gl::pow(gl::pow(beta[s], n1), k2)));
    }
}

// Synthetic uniform inputs: counter-based SplitMix64 finaliser of seed + (i+1)*gamma.
__global__ void k_fill_uniform(u64* d, u64 count, u64 seed) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    u64 z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    d[i] = z >= gl::P ? z - gl::P : z;
}

int launch_fill_uniform(u64* d, u64 count, u64 seed, hipStream_t st) {
    if (!count) return TVH_OK;
    hipLaunchKernelGGL(k_fill_uniform, dim3((u32)((count + 255) / 256)), dim3(256), 0, st, d,
                       count, seed);
    return check_launch();
}

// XFE interleaved <-> limb-major.  Interleaved: (c*len + r)*3 + k; limb-major: (3c + k)*len + r.
__global__ void k_xfe_relayout(const u64* __restrict__ src, u64* __restrict__ dst, u64 cols,
                               u64 len, int to_limb_major) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cols * len * 3) return;
    // i indexes the interleaved position
    u64 k = i % 3, cr = i / 3, r = cr % len, c = cr / len;
    u64 lm = (3 * c + k) * len + r;
    if (to_limb_major) dst[lm] = src[i];
    else dst[i] = src[lm];
}

// ---------------------------------------------------------------- host side

// Tuned K2 form per (log2 N, log2 E); 0 = not set.
static std::atomic<int> g_variant[kMaxLogN + 1][5];

int lde_variant(u32 logN, u32 logE) {
    int v = (logN <= kMaxLogN && logE <= 4) ? g_variant[logN][logE].load() : 0;
    if (v) return v;
    // default (measured on MI355X at widths 156..379, 2^12..2^21 rows): the split
    // form with 2^12-element row tiles for 2^13..2^16 rows, 2^13-element row
    // tiles above (column tiles 2^12, 2^13 from 2^19 rows), 2^13 for the inverse
    // rows only at 2^12 rows, the single fused kernel below
    if (logN >= 19) return LDE_SPLIT | (1 << 2) | LDE_ROW_HALF;
    if (logN >= 17) return LDE_SPLIT | (2 << 2) | LDE_ROW_HALF;
    if (logN >= 13) return LDE_SPLIT | (2 << 2) | LDE_ROW_QUARTER;
    if (logN == 12) return LDE_SPLIT | (1 << 2) | LDE_INV_HALF;
    return LDE_WIDE | (2 << 2);
}

bool lde_variant_valid(int v) {
    int form = v & 3, tile = (v >> 2) & 3, half = (v >> 4) & 1, ihalf = (v >> 5) & 1,
        quarter = (v >> 6) & 1;
    return v == LDE_AUTO || (v < 128 && (form == LDE_WIDE || form == LDE_SPLIT) && tile <= 2 &&
                             half + ihalf + quarter <= (form == LDE_SPLIT));
}

int set_lde_variant(u32 logN, u32 logE, int v) {
    if (logN > kMaxLogN || logE > 4 || !lde_variant_valid(v))
        return set_error(TVH_ERR_INVALID_ARG, "lde variant (%u, %u, %d)", logN, logE, v);
    g_variant[logN][logE].store(v);
    return TVH_OK;
}

static u32 choose_split(u32 logN, u32 E, u32 row_lt) {
    // N2 (the contiguous row step) as long as one row tile holds all E cosets
    // of a row (E * N2 <= 2^row_lt, N2 <= 4096: the expansion overwrites its
    // input row in place, so a row's cosets cannot be split over blocks);
    // N1 = N / N2 takes the strided step, whose HBM segments are (column
    // tile) / N1 elements wide.
    u32 logE = log2u(E);
    u32 bmax = row_lt - logE;
    if (bmax > 12) bmax = 12;
    u32 b = logN < bmax ? logN : bmax;
    return logN - b;
}

static u32 tile_log_nt(u32 logL, u32 lt = LOG_TILE) {
    // transforms per block of a 2^lt tile: 2^lt / max(L, 32)
    u32 l = logL < 5 ? 5 : logL;
    return lt - l;
}

static size_t lds_bytes(u32 logL, u32 logNT) {
    return ((size_t)padded_len(1 << logL) << logNT) * sizeof(u64);
}

struct ConstKey {
    u32 logN, a, E;
    u64 toff, goff;
    int dev;
    bool operator<(const ConstKey& o) const {
        return std::tie(logN, a, E, toff, goff, dev) <
               std::tie(o.logN, o.a, o.E, o.toff, o.goff, o.dev);
    }
};
struct ConstVal {
    u64* pa;
    u64* pb;
};
static std::mutex g_const_mu;
static std::map<ConstKey, ConstVal> g_consts;
// At most this many tables (each <= E*(N1+N2)*8 <= 1 MB) are kept; the cache
// never evicts, since kernels in flight may read any entry.
constexpr size_t kMaxConstEntries = 256;

// The coset tables of one (shape, split, offsets) key: found in the cache, or
// computed now (one allocation and a wait on a private stream; tvh_lde_prepare
// does this ahead of time).
static int lde_consts(DeviceState* ds, u32 logN, u32 a, u32 E, u64 toff, u64 goff, ConstVal* out) {
    ConstKey key{logN, a, E, toff, goff, ds->device};
    std::lock_guard<std::mutex> lk(g_const_mu);
    auto it = g_consts.find(key);
    if (it != g_consts.end()) {
        *out = it->second;
        return TVH_OK;
    }
    if (g_consts.size() >= kMaxConstEntries)
        return set_error(TVH_ERR_RANGE,
                         "more than %zu distinct LDE shape/offset tables; release them with "
                         "tvh_lde_release_constants",
                         kMaxConstEntries);
    u32 b = logN - a;
    u64 n1 = 1ull << a, n2 = 1ull << b;
    u64 M = (u64)E << logN;
    u64 wM = gl::root_of_unity(log2u(M));
    u64 oinv = gl::inv(toff);
    u64 beta[16];
    u64 ws = 1;
    for (u32 s = 0; s < E; s++) {
        beta[s] = gl::mul(gl::mul(goff, ws), oinv);
        ws = gl::mul(ws, wM);
    }
    u64 ninv = gl::inv(1ull << logN);
    u64* mem = nullptr;
    size_t bytes = (E * (n1 + n2) + 16) * sizeof(u64);
    if (hipMalloc(&mem, bytes) != hipSuccess) return set_error(TVH_ERR_OOM, "lde constants");
    u64* dbeta = mem + E * (n1 + n2);
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
        return set_error(TVH_ERR_DEVICE, "stream create");
    int rc = TVH_OK;
    if (hipMemcpyAsync(dbeta, beta, sizeof(beta), hipMemcpyHostToDevice, st) != hipSuccess)
        rc = set_error(TVH_ERR_DEVICE, "copy beta");
    if (!rc) {
        u64 total = E * (n1 + n2);
        hipLaunchKernelGGL(k_lde_consts, dim3((u32)((total + 255) / 256)), dim3(256), 0, st, mem,
                           mem + E * n1, a, b, E, dbeta, ninv);
        rc = check_launch();
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = set_error(TVH_ERR_DEVICE, "sync");
    (void)hipStreamDestroy(st);
    if (rc) {
        (void)hipFree(mem);
        return rc;
    }
    ConstVal v{mem, mem + E * n1};
    g_consts[key] = v;
    *out = v;
    return TVH_OK;
}

int lde_prepare_consts(u64 n, u64 m, u64 toff, u64 goff) {
    if (!is_pow2(n) || !is_pow2(m)) return set_error(TVH_ERR_NOT_POW2, "lde lengths");
    if (m < n) return set_error(TVH_ERR_INVALID_ARG, "target_len < trace_len");
    if (m / n > 16) return set_error(TVH_ERR_UNSUPPORTED, "expansion > 16");
    if (toff == 0 || toff >= gl::P || goff >= gl::P)
        return set_error(TVH_ERR_INVALID_ARG, "offsets must be canonical, trace offset nonzero");
    u32 logN = log2u(n);
    if (logN > 23) return set_error(TVH_ERR_UNSUPPORTED, "trace_len > 2^23");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    u32 E = (u32)(m / n);
    // every row tile a kernel variant can use (2^12..2^14 elements) gives its own split
    for (u32 lt = LOG_TILE - 2; lt <= LOG_TILE; lt++) {
        u32 a = choose_split(logN, E, lt);
        if (a > 12) continue;
        ConstVal cv;
        int rc = lde_consts(ds, logN, a, E, toff, goff, &cv);
        if (rc) return rc;
    }
    return TVH_OK;
}

int lde_release_consts() {
    std::lock_guard<std::mutex> lk(g_const_mu);
    if (hipDeviceSynchronize() != hipSuccess) return set_error(TVH_ERR_DEVICE, "device sync");
    for (auto& kv : g_consts) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        (void)hipSetDevice(kv.first.dev);
        (void)hipFree(kv.second.pa);
        (void)hipSetDevice(dev);
    }
    g_consts.clear();
    return TVH_OK;
}

int prepare_lde(const u64* trace, u64 cols, u64 n, const RandSpec& rs, u64 toff, u64 goff,
                u64 m, bool truncate, u64* out, LdePlan* P, bool batch) {
    P->k1 = P->k3 = P->split = false;
    P->k2_variant = lde_variant(log2u(n), log2u(m / (n ? n : 1)));
    // multi-table launches run the column and row kernels with one tile each
    // (kBatchColLt, kBatchRowLt) for every table of the launch
    P->col_lt = batch ? kBatchColLt : LOG_TILE - ((P->k2_variant >> 2) & 3);
    P->row_lt = batch                                  ? kBatchRowLt
                : (P->k2_variant & LDE_ROW_QUARTER) ? LOG_TILE - 2
                : (P->k2_variant & LDE_ROW_HALF)    ? LOG_TILE - 1
                                                     : LOG_TILE;
    P->inv_lt = P->row_lt < LOG_TILE                 ? P->row_lt
                : (P->k2_variant & LDE_INV_HALF)    ? LOG_TILE - 1
                                                     : LOG_TILE;
    if ((P->k2_variant & 3) != LDE_SPLIT) P->row_lt = LOG_TILE;  // the wide kernel's tile
    P->s1_blocks = P->s2_blocks = P->s3_blocks = 0;
    if (cols == 0) return TVH_OK;
    if (!is_pow2(n) || !is_pow2(m)) return set_error(TVH_ERR_NOT_POW2, "lde lengths");
    if (m < n) return set_error(TVH_ERR_INVALID_ARG, "target_len < trace_len");
    u64 E64 = m / n;
    if (E64 > 16) return set_error(TVH_ERR_UNSUPPORTED, "expansion > 16");
    if (rs.k > n) return set_error(TVH_ERR_UNSUPPORTED, "rand_len > trace_len");
    if (toff == 0 || toff >= gl::P || goff >= gl::P)
        return set_error(TVH_ERR_INVALID_ARG, "offsets must be canonical, trace offset nonzero");
    u32 logN = log2u(n), E = (u32)E64, logE = log2u(E64);
    if (logN > 23) return set_error(TVH_ERR_UNSUPPORTED, "trace_len > 2^23");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    u32 a = choose_split(logN, E, P->row_lt);
    u32 b = logN - a;
    if (a > 12) return set_error(TVH_ERR_UNSUPPORTED, "split");
    ConstVal cv;
    int rc = lde_consts(ds, logN, a, E, toff, goff, &cv);
    if (rc) return rc;

    // K1: INTT step 1 (transforms (c, n2)); a size-1 step is the identity
    // map trace -> Y, so K2 then reads the trace itself.
    P->k1 = a > 0;
    if (P->k1) {
        u32 logNT = tile_log_nt(a, P->col_lt);
        u64 ntr = cols << b;
        P->s1 = Step1Args{trace, n, out, m, logN, a, E, ntr, logNT};
        P->s1_blocks = (u32)((ntr + (1ull << logNT) - 1) >> logNT);
        P->s1_lds = (size_t)fused_pitch(1 << a, 1 << logNT, true) * (1u << logNT) * sizeof(u64);
    }
    // K2: INTT step 2 + coset scaling + forward step 1
    {
        LdeArgs& A = P->s2;
        A.out = out;
        A.logN = logN;
        A.a = a;
        A.E = E;
        A.logE = logE;
        A.cols = cols;
        A.M = m;
        A.pa = cv.pa;
        A.pb = cv.pb;
        A.toff = toff;
        A.rs = rs;
        u64 gN = gl::pow(goff, n), oN = gl::pow(toff, n);
        u64 wE = gl::root_of_unity(logE), ws = 1;
        u64 nmod = n % gl::P;
        for (u32 s = 0; s < 16; s++) {
            u64 d = (truncate && E == 1) ? gl::neg(oN) : gl::sub(gl::mul(gN, ws), oN);
            A.delta_n[s] = gl::mul(d, nmod);
            ws = gl::mul(ws, wE);
        }
        u32 logNT = P->row_lt - (b + logE);  // NT * E * N2 = 2^row_lt
        if (b < 5) logNT = P->row_lt - 5 - logE;
        u64 nrows = cols << a;
        if ((1ull << logNT) > nrows) logNT = log2u(nrows);
        A.logNT = logNT;
        A.src = P->k1 ? out : trace;
        A.src_col_stride = P->k1 ? m : n;
        A.src_row_stride = P->k1 ? ((u64)E << b) : 0;
        P->s2_blocks = (u32)((nrows + (1ull << logNT) - 1) >> logNT);
        P->s2_lds = lds_bytes(b, logNT + logE);
        if ((P->k2_variant & 3) == LDE_SPLIT)  // k_lde_expand: transform-fastest pitch
            P->s2_lds = (size_t)fused_pitch(1 << b, (int)(E << logNT), true) * (E << logNT) * sizeof(u64);
        P->split = (P->k2_variant & 3) == LDE_SPLIT;
        if (P->split) {  // inverse rows first (TILE / N2 rows per block), then expansion from out
            u32 li = tile_log_nt(b, P->inv_lt);
            if ((1ull << li) > nrows) li = log2u(nrows);
            P->inv_logNT = li;
            P->inv_blocks = (u32)((nrows + (1ull << li) - 1) >> li);
            P->inv_lds = lds_bytes(b, li);
            P->s2inv = A;
            P->s2inv.logNT = li;
            A.src = out;
            A.src_col_stride = m;
            A.src_row_stride = (u64)E << b;
        }
    }
    // K3: forward step 2 for all cosets (identity when N1 = 1)
    P->k3 = a > 0;
    if (P->k3) {
        u32 logNT = tile_log_nt(a, P->col_lt);
        u64 ntr = cols << (b + logE);
        P->s3 = Step3Args{out, logN, a, logE, cols, m, logNT};
        P->s3_blocks = (u32)((ntr + (1ull << logNT) - 1) >> logNT);
        P->s3_lds = (size_t)fused_pitch(1 << a, 1 << logNT, true) * (1u << logNT) * sizeof(u64);
    }
    return TVH_OK;
}

int launch_lde(const u64* trace, u64 cols, u64 n, const RandSpec& rs, u64 toff, u64 goff, u64 m,
               bool truncate, u64* out, hipStream_t st) {
    LdePlan P;
    int rc = prepare_lde(trace, cols, n, rs, toff, goff, m, truncate, out, &P, false);
    if (rc || cols == 0) return rc;
    DeviceState* ds = current_device();
    const dim3 ctpb(1u << (P.col_lt - 4));
    if (P.k1) {
        if (P.col_lt == 12)
            hipLaunchKernelGGL((k_step1<true, 12>), dim3(P.s1_blocks), ctpb, P.s1_lds, st, P.s1,
                               ds->roots);
        else if (P.col_lt == 13)
            hipLaunchKernelGGL((k_step1<true, 13>), dim3(P.s1_blocks), ctpb, P.s1_lds, st, P.s1,
                               ds->roots);
        else
            hipLaunchKernelGGL((k_step1<true, 14>), dim3(P.s1_blocks), ctpb, P.s1_lds, st, P.s1,
                               ds->roots);
        if ((rc = check_launch())) return rc;
    }
    if (P.split) {
        const dim3 rtpb(1u << (P.row_lt - 4)), itpb(1u << (P.inv_lt - 4));
        if (P.inv_lt == LOG_TILE)
            hipLaunchKernelGGL(k_lde_rowinv<LOG_TILE>, dim3(P.inv_blocks), itpb, P.inv_lds, st,
                               P.s2inv, ds->roots);
        else if (P.inv_lt == LOG_TILE - 1)
            hipLaunchKernelGGL(k_lde_rowinv<LOG_TILE - 1>, dim3(P.inv_blocks), itpb, P.inv_lds, st,
                               P.s2inv, ds->roots);
        else
            hipLaunchKernelGGL(k_lde_rowinv<LOG_TILE - 2>, dim3(P.inv_blocks), itpb, P.inv_lds, st,
                               P.s2inv, ds->roots);
        if ((rc = check_launch())) return rc;
        if (P.row_lt == LOG_TILE)
            hipLaunchKernelGGL(k_lde_expand<LOG_TILE>, dim3(P.s2_blocks), rtpb, P.s2_lds, st, P.s2,
                               ds->roots);
        else if (P.row_lt == LOG_TILE - 1)
            hipLaunchKernelGGL(k_lde_expand<LOG_TILE - 1>, dim3(P.s2_blocks), rtpb, P.s2_lds, st,
                               P.s2, ds->roots);
        else
            hipLaunchKernelGGL(k_lde_expand<LOG_TILE - 2>, dim3(P.s2_blocks), rtpb, P.s2_lds, st,
                               P.s2, ds->roots);
    } else
        hipLaunchKernelGGL(k_lde_step2, dim3(P.s2_blocks), dim3(TPB), P.s2_lds, st, P.s2,
                           ds->roots);
    if ((rc = check_launch())) return rc;
    if (P.k3) {
        if (P.col_lt == 12)
            hipLaunchKernelGGL(k_lde_step3<12>, dim3(P.s3_blocks), ctpb, P.s3_lds, st, P.s3,
                               ds->roots);
        else if (P.col_lt == 13)
            hipLaunchKernelGGL(k_lde_step3<13>, dim3(P.s3_blocks), ctpb, P.s3_lds, st, P.s3,
                               ds->roots);
        else
            hipLaunchKernelGGL(k_lde_step3<14>, dim3(P.s3_blocks), ctpb, P.s3_lds, st, P.s3,
                               ds->roots);
        if ((rc = check_launch())) return rc;
    }
    return TVH_OK;
}

int launch_lde_multi(int phase, const void* d_args, const u32* d_start, u32 nseg, u32 blocks,
                     size_t lds, hipStream_t st) {
    if (blocks == 0) return TVH_OK;
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    const dim3 ctpb(1u << (kBatchColLt - 4)), rtpb(1u << (kBatchRowLt - 4));
    switch (phase) {
        case LDE_PHASE_STEP1:
            hipLaunchKernelGGL((k_step1_multi<true, kBatchColLt>), dim3(blocks), ctpb, lds, st,
                               (const Step1Args*)d_args, d_start, nseg, ds->roots);
            break;
        case LDE_PHASE_ROWINV:
            hipLaunchKernelGGL(k_lde_rowinv_multi<kBatchRowLt>, dim3(blocks), rtpb, lds, st,
                               (const LdeArgs*)d_args, d_start, nseg, ds->roots);
            break;
        case LDE_PHASE_STEP2:
            hipLaunchKernelGGL(k_lde_step2_multi, dim3(blocks), dim3(TPB), lds, st,
                               (const LdeArgs*)d_args, d_start, nseg, ds->roots);
            break;
        case LDE_PHASE_EXPAND:
            hipLaunchKernelGGL(k_lde_expand_multi<kBatchRowLt>, dim3(blocks), rtpb, lds, st,
                               (const LdeArgs*)d_args, d_start, nseg, ds->roots);
            break;
        default:
            hipLaunchKernelGGL(k_lde_step3_multi<kBatchColLt>, dim3(blocks), ctpb, lds, st,
                               (const Step3Args*)d_args, d_start, nseg, ds->roots);
    }
    return check_launch();
}

u64 ntt_scratch_elems(u64 n, u64 cols) {
    // columns go through the scratch in groups of at most 256 MB
    u64 group = cols;
    u64 max_group = (256ull << 20) / (n * sizeof(u64));
    if (max_group < 1) max_group = 1;
    if (group > max_group) group = max_group;
    return group * n;
}

int launch_ntt_batched(u64* data, u64 n, u64 cols, bool inverse, u64* scratch, u64 scratch_elems,
                       hipStream_t st) {
    if (cols == 0) return TVH_OK;
    if (!is_pow2(n)) return set_error(TVH_ERR_NOT_POW2, "NTT size must be a power of 2");
    u32 logN = log2u(n);
    if (logN > kMaxLogN) return set_error(TVH_ERR_UNSUPPORTED, "NTT size > 2^24");
    DeviceState* ds = current_device();
    if (!ds) return TVH_ERR_DEVICE;
    u32 a = logN / 2, b = logN - a;
    if (b > 12) {
        b = 12;
        a = logN - b;
    }
    // step 1 writes each group of columns to the scratch, step 2 writes it back
    u64 group = scratch_elems / n;
    if (!scratch || group == 0)
        return set_error(TVH_ERR_INVALID_ARG, "NTT scratch of %llu elements < n = %llu",
                         (unsigned long long)scratch_elems, (unsigned long long)n);
    if (group > cols) group = cols;
    u64 ninv = gl::inv(n % gl::P);
    int rc;
    for (u64 c0 = 0; c0 < cols; c0 += group) {
        u64 g = cols - c0 < group ? cols - c0 : group;
        u64* d = data + c0 * n;
        {
            u32 logNT = tile_log_nt(a);
            u64 ntr = g << b;
            u64 blocks = (ntr + (1ull << logNT) - 1) >> logNT;
            Step1Args A{d, n, scratch, n, logN, a, 1u, ntr, logNT};
            const size_t s1_lds =
                (size_t)fused_pitch(1 << a, 1 << logNT, true) * (1u << logNT) * sizeof(u64);
            if (inverse)
                hipLaunchKernelGGL(k_step1<true>, dim3((u32)blocks), dim3(TPB), s1_lds, st, A,
                                   ds->roots);
            else
                hipLaunchKernelGGL(k_step1<false>, dim3((u32)blocks), dim3(TPB), s1_lds, st, A,
                                   ds->roots);
            if ((rc = check_launch())) return rc;
        }
        {
            u32 logNT = tile_log_nt(b);
            u64 nrows = g << a;
            if ((1ull << logNT) > nrows) logNT = log2u(nrows);
            u64 blocks = (nrows + (1ull << logNT) - 1) >> logNT;
            if (inverse)
                hipLaunchKernelGGL(k_ntt_step2<true>, dim3((u32)blocks), dim3(TPB),
                                   lds_bytes(b, logNT), st, scratch, d, logN, a, nrows, logNT,
                                   ninv, ds->roots);
            else
                hipLaunchKernelGGL(k_ntt_step2<false>, dim3((u32)blocks), dim3(TPB),
                                   lds_bytes(b, logNT), st, scratch, d, logN, a, nrows, logNT,
                                   ninv, ds->roots);
            if ((rc = check_launch())) return rc;
        }
    }
    return TVH_OK;
}

int ntt_set_attributes() {
    const int bytes = 150 * 1024;
    const void* fns[] = {(const void*)k_step1<true>, (const void*)k_step1<false>,
                         (const void*)k_step1<true, 13>, (const void*)k_step1<true, 12>,
                         (const void*)k_ntt_step2<true>, (const void*)k_ntt_step2<false>,
                         (const void*)k_lde_step2, (const void*)k_lde_step3<14>,
                         (const void*)k_lde_step3<13>, (const void*)k_lde_step3<12>,
                         (const void*)k_step1_multi<true, kBatchColLt>,
                         (const void*)k_lde_step2_multi,
                         (const void*)k_lde_step3_multi<kBatchColLt>, (const void*)k_lde_rowinv<LOG_TILE>,
                         (const void*)k_lde_expand<LOG_TILE>, (const void*)k_lde_rowinv<LOG_TILE - 1>,
                         (const void*)k_lde_expand<LOG_TILE - 1>, (const void*)k_lde_rowinv<LOG_TILE - 2>,
                         (const void*)k_lde_expand<LOG_TILE - 2>, (const void*)k_lde_rowinv_multi<kBatchRowLt>,
                         (const void*)k_lde_expand_multi<kBatchRowLt>,
};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
            return set_error(TVH_ERR_DEVICE, "LDS attribute");
    return TVH_OK;
}

int launch_xfe_interleave(const u64* src, u64* dst, u64 cols, u64 len, bool to_limb_major,
                          hipStream_t st) {
    u64 total = cols * len * 3;
    if (!total) return TVH_OK;
    hipLaunchKernelGGL(k_xfe_relayout, dim3((u32)((total + 255) / 256)), dim3(256), 0, st, src,
                       dst, cols, len, (int)to_limb_major);
    return check_launch();
}

}  // namespace tvh

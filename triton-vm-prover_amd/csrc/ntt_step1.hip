// Column (strided) step of the batched NTT and of the LDE's inverse transform,
// and the plain NTT's row step: kernels and their launchers.
#include "ntt_dev.hpp"

namespace tvh {

template <bool INV, int LT>
__global__ void __launch_bounds__(1 << (LT - 4), 4) k_step1(Step1Args A, Roots rt) {
    step1_body<INV, LT>(A, blockIdx.x, rt);
}

template <bool INV, int LT>
__global__ void __launch_bounds__(1 << (LT - 4), 4) k_step1_multi(const Step1Args* __restrict__ args,
                                                                  const u32* __restrict__ start,
                                                                  u32 nseg, Roots rt) {
    u32 seg = find_segment(start, nseg, blockIdx.x);
    const Step1Args& A = args[seg];
    step1_body<INV, LT>(A, blockIdx.x - start[seg], rt);
}

template <int LT>
__global__ void __launch_bounds__(1 << (LT - 4), 4) k_step1_scaled(Step1Args A, NttScaleDev S,
                                                                   Roots rt) {
    step1_scaled_body<LT>(A, S, blockIdx.x, rt);
}

// Plain NTT row step.  Rows s = (c, k1) of N2 = 2^b elements at src[s*N2 + n2]
// (the column step's output); DFT over n2 -> k2, times N^-1 when INV; element
// k2 of row (c, k1) to dst[c*N + k1 + N1*k2].  Fused endpoints: the first pass
// reads the rows from HBM element-fastest, the last pass writes
// transform-fastest, so adjacent lanes write adjacent k1 (runs as long as the
// block's row count).  Offsets stay below N <= 2^24 plus one tile.
struct PlainRowLoad {
    const u64* src;
    u64 s0, nrows;
    u32 b;
    template <int R>
    __device__ __forceinline__ void load(int t, int j, int per, u64 (&v)[R]) const {
        if (s0 + t < nrows) {
            const u64* p = src + (s0 << b);
            const u32 o = (u32)t << b;
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = p[o + (u32)(j + r * per)];
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = 0;
        }
    }
};
template <bool INV>
struct PlainRowStore {
    u64* dst;
    u64 n, s0, nrows, ninv_m;  // ninv_m: N^-1 in Montgomery form
    u32 a;
    template <int R>
    __device__ __forceinline__ void store(int t, int j, int Ns, const u64 (&v)[R]) const {
        u64 s = s0 + t;
        if (s >= nrows) return;
        const u64 mask = (1ull << a) - 1;
        u64* p = dst + (s0 >> a) * n + (s0 & mask);
        const u32 o = (u32)(((s >> a) - (s0 >> a)) * n + (s & mask) - (s0 & mask));
        int k = j & (Ns - 1), base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) {
            // (canonical in, canonical out: the product is below p^2)
            u64 x = INV ? gl::mont_mul(v[r], ninv_m) : v[r];
            p[o + ((u32)(base + r * Ns) << a)] = x;
        }
    }
};

template <bool INV, int LT>
__global__ void __launch_bounds__(1 << (LT - 4)) k_ntt_step2(const u64* __restrict__ src,
                                                             u64* __restrict__ dst, u32 logN, u32 a,
                                                             u64 nrows, u32 logNT, u64 ninv_m,
                                                             Roots rt) {
    extern __shared__ u64 lds[];
    const u32 b = logN - a;
    const u64 s0 = (u64)blockIdx.x << logNT;
    const PlainRowLoad ld{src, s0, nrows, b};
    const PlainRowStore<INV> st{dst, 1ull << logN, s0, nrows, ninv_m, a};
    fused_dft<INV, (1 << (LT - 4)), (1 << LT), false, true>(lds, b, 1 << logNT, logNT,
                                                           rt.r4096[INV], ld, st);
}

int launch_k_step1(bool inv, u32 lt, u32 blocks, size_t lds, hipStream_t st, const Step1Args& A,
                   const Roots& rt) {
    const dim3 tpb(1u << (lt - 4));
    if (!inv) {
        if (lt == LOG_TILE - 1)
            hipLaunchKernelGGL((k_step1<false, LOG_TILE - 1>), dim3(blocks), tpb, lds, st, A, rt);
        else if (lt == LOG_TILE)
            hipLaunchKernelGGL((k_step1<false, LOG_TILE>), dim3(blocks), tpb, lds, st, A, rt);
        else
            return set_error(TVH_ERR_UNSUPPORTED, "forward step 1 tile %u", lt);
    } else if (lt == 12) {
        hipLaunchKernelGGL((k_step1<true, 12>), dim3(blocks), tpb, lds, st, A, rt);
    } else if (lt == 13) {
        hipLaunchKernelGGL((k_step1<true, 13>), dim3(blocks), tpb, lds, st, A, rt);
    } else {
        hipLaunchKernelGGL((k_step1<true, 14>), dim3(blocks), tpb, lds, st, A, rt);
    }
    return check_launch();
}

int launch_k_step1_scaled(u32 lt, u32 blocks, size_t lds, hipStream_t st, const Step1Args& A,
                          const NttScaleDev& S, const Roots& rt) {
    const dim3 tpb(1u << (lt - 4));
    if (lt == LOG_TILE - 1)
        hipLaunchKernelGGL((k_step1_scaled<LOG_TILE - 1>), dim3(blocks), tpb, lds, st, A, S, rt);
    else if (lt == LOG_TILE)
        hipLaunchKernelGGL((k_step1_scaled<LOG_TILE>), dim3(blocks), tpb, lds, st, A, S, rt);
    else
        return set_error(TVH_ERR_UNSUPPORTED, "scaled step 1 tile %u", lt);
    return check_launch();
}

int launch_k_step1_multi(const Step1Args* args, const u32* start, u32 nseg, u32 blocks, size_t lds,
                         hipStream_t st, const Roots& rt) {
    hipLaunchKernelGGL((k_step1_multi<true, kBatchColLt>), dim3(blocks), dim3(1u << (kBatchColLt - 4)),
                       lds, st, args, start, nseg, rt);
    return check_launch();
}

int launch_k_ntt_step2(bool inv, u32 lt, u32 blocks, size_t lds, hipStream_t st, const u64* src,
                       u64* dst, u32 logN, u32 a, u64 nrows, u32 logNT, u64 ninv_m,
                       const Roots& rt) {
    const dim3 tpb(1u << (lt - 4));
#define TVH_STEP2(I, L) \
    hipLaunchKernelGGL((k_ntt_step2<I, L>), dim3(blocks), tpb, lds, st, src, dst, logN, a, nrows, \
                       logNT, ninv_m, rt)
    if (lt == LOG_TILE) {
        if (inv) TVH_STEP2(true, LOG_TILE);
        else TVH_STEP2(false, LOG_TILE);
    } else if (lt == LOG_TILE - 1) {
        if (inv) TVH_STEP2(true, LOG_TILE - 1);
        else TVH_STEP2(false, LOG_TILE - 1);
    } else {
        return set_error(TVH_ERR_UNSUPPORTED, "row step tile %u", lt);
    }
#undef TVH_STEP2
    return check_launch();
}

int set_attributes_step1(int bytes) {
    const void* fns[] = {(const void*)k_step1<true, 14>, (const void*)k_step1<false, 14>,
                         (const void*)k_step1<true, 13>, (const void*)k_step1<true, 12>,
                         (const void*)k_step1<false, 13>, (const void*)k_step1_scaled<13>,
                         (const void*)k_step1_scaled<14>,
                         (const void*)k_step1_multi<true, kBatchColLt>,
                         (const void*)k_ntt_step2<true, 14>, (const void*)k_ntt_step2<false, 14>,
                         (const void*)k_ntt_step2<true, 13>, (const void*)k_ntt_step2<false, 13>};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
            return set_error(TVH_ERR_DEVICE, "LDS attribute");
    return TVH_OK;
}

}  // namespace tvh

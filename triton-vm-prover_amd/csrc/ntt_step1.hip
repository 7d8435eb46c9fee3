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

int launch_k_step1(bool inv, u32 lt, u32 blocks, size_t lds, hipStream_t st, const Step1Args& A,
                   const Roots& rt) {
    const dim3 tpb(1u << (lt - 4));
    if (!inv) {
        if (lt != LOG_TILE) return set_error(TVH_ERR_UNSUPPORTED, "forward step 1 tile %u", lt);
        hipLaunchKernelGGL((k_step1<false, LOG_TILE>), dim3(blocks), tpb, lds, st, A, rt);
    } else if (lt == 12) {
        hipLaunchKernelGGL((k_step1<true, 12>), dim3(blocks), tpb, lds, st, A, rt);
    } else if (lt == 13) {
        hipLaunchKernelGGL((k_step1<true, 13>), dim3(blocks), tpb, lds, st, A, rt);
    } else {
        hipLaunchKernelGGL((k_step1<true, 14>), dim3(blocks), tpb, lds, st, A, rt);
    }
    return check_launch();
}

int launch_k_step1_multi(const Step1Args* args, const u32* start, u32 nseg, u32 blocks, size_t lds,
                         hipStream_t st, const Roots& rt) {
    hipLaunchKernelGGL((k_step1_multi<true, kBatchColLt>), dim3(blocks), dim3(1u << (kBatchColLt - 4)),
                       lds, st, args, start, nseg, rt);
    return check_launch();
}

int launch_k_ntt_step2(bool inv, u32 blocks, size_t lds, hipStream_t st, const u64* src, u64* dst,
                       u32 logN, u32 a, u64 nrows, u32 logNT, u64 ninv, const Roots& rt) {
    if (inv)
        hipLaunchKernelGGL(k_ntt_step2<true>, dim3(blocks), dim3(TPB), lds, st, src, dst, logN, a,
                           nrows, logNT, ninv, rt);
    else
        hipLaunchKernelGGL(k_ntt_step2<false>, dim3(blocks), dim3(TPB), lds, st, src, dst, logN, a,
                           nrows, logNT, ninv, rt);
    return check_launch();
}

int set_attributes_step1(int bytes) {
    const void* fns[] = {(const void*)k_step1<true, 14>, (const void*)k_step1<false, 14>,
                         (const void*)k_step1<true, 13>, (const void*)k_step1<true, 12>,
                         (const void*)k_step1_multi<true, kBatchColLt>,
                         (const void*)k_ntt_step2<true>, (const void*)k_ntt_step2<false>};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
            return set_error(TVH_ERR_DEVICE, "LDS attribute");
    return TVH_OK;
}

}  // namespace tvh

// Row steps of the LDE: the inverse row DFTs of the split form (k_lde_rowinv)
// and the single-kernel wide form (k_lde_step2): kernels and their launchers.
#include "ntt_dev.hpp"

namespace tvh {

__global__ void __launch_bounds__(TPB) k_lde_step2(LdeArgs A, Roots rt) {
    lde_step2_body<true>(A, blockIdx.x, rt);
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

__global__ void __launch_bounds__(TPB) k_lde_step2_multi(const LdeArgs* __restrict__ args,
                                                         const u32* __restrict__ start, u32 nseg,
                                                         Roots rt) {
    u32 seg = find_segment(start, nseg, blockIdx.x);
    const LdeArgs& A = args[seg];  // read in place: a private copy would spill delta_n
    lde_step2_body<true>(A, blockIdx.x - start[seg], rt);
}

int launch_k_lde_rowinv(u32 lt, u32 blocks, size_t lds, hipStream_t st, const LdeArgs& A,
                        const Roots& rt) {
    const dim3 tpb(1u << (lt - 4));
    if (lt == LOG_TILE)
        hipLaunchKernelGGL(k_lde_rowinv<LOG_TILE>, dim3(blocks), tpb, lds, st, A, rt);
    else if (lt == LOG_TILE - 1)
        hipLaunchKernelGGL(k_lde_rowinv<LOG_TILE - 1>, dim3(blocks), tpb, lds, st, A, rt);
    else
        hipLaunchKernelGGL(k_lde_rowinv<LOG_TILE - 2>, dim3(blocks), tpb, lds, st, A, rt);
    return check_launch();
}

int launch_k_lde_rowinv_multi(const LdeArgs* args, const u32* start, u32 nseg, u32 blocks,
                              size_t lds, hipStream_t st, const Roots& rt) {
    hipLaunchKernelGGL(k_lde_rowinv_multi<kBatchRowLt>, dim3(blocks), dim3(1u << (kBatchRowLt - 4)),
                       lds, st, args, start, nseg, rt);
    return check_launch();
}

int launch_k_lde_step2(u32 blocks, size_t lds, hipStream_t st, const LdeArgs& A, const Roots& rt) {
    hipLaunchKernelGGL(k_lde_step2, dim3(blocks), dim3(TPB), lds, st, A, rt);
    return check_launch();
}

int launch_k_lde_step2_multi(const LdeArgs* args, const u32* start, u32 nseg, u32 blocks,
                             size_t lds, hipStream_t st, const Roots& rt) {
    hipLaunchKernelGGL(k_lde_step2_multi, dim3(blocks), dim3(TPB), lds, st, args, start, nseg, rt);
    return check_launch();
}

int set_attributes_rows(int bytes) {
    const void* fns[] = {(const void*)k_lde_step2, (const void*)k_lde_step2_multi,
                         (const void*)k_lde_rowinv<LOG_TILE>, (const void*)k_lde_rowinv<LOG_TILE - 1>,
                         (const void*)k_lde_rowinv<LOG_TILE - 2>,
                         (const void*)k_lde_rowinv_multi<kBatchRowLt>};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
            return set_error(TVH_ERR_DEVICE, "LDS attribute");
    return TVH_OK;
}

}  // namespace tvh

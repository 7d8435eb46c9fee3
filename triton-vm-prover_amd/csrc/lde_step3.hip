// Column step of the LDE's forward transforms (k_lde_step3): kernels and launchers.
#include "ntt_dev.hpp"

namespace tvh {

template <int LT>
__global__ void __launch_bounds__(1 << (LT - 4), 4) k_lde_step3(Step3Args A, Roots rt) {
    lde_step3_body<LT>(A, blockIdx.x, rt);
}

template <int LT>
__global__ void __launch_bounds__(1 << (LT - 4), 4) k_lde_step3_multi(const Step3Args* __restrict__ args,
                                                                      const u32* __restrict__ start,
                                                                      u32 nseg, Roots rt) {
    u32 seg = find_segment(start, nseg, blockIdx.x);
    const Step3Args& A = args[seg];
    lde_step3_body<LT>(A, blockIdx.x - start[seg], rt);
}

int launch_k_lde_step3(u32 lt, u32 blocks, size_t lds, hipStream_t st, const Step3Args& A,
                       const Roots& rt) {
    const dim3 tpb(1u << (lt - 4));
    if (lt == 12)
        hipLaunchKernelGGL(k_lde_step3<12>, dim3(blocks), tpb, lds, st, A, rt);
    else if (lt == 13)
        hipLaunchKernelGGL(k_lde_step3<13>, dim3(blocks), tpb, lds, st, A, rt);
    else
        hipLaunchKernelGGL(k_lde_step3<14>, dim3(blocks), tpb, lds, st, A, rt);
    return check_launch();
}

int launch_k_lde_step3_multi(const Step3Args* args, const u32* start, u32 nseg, u32 blocks,
                             size_t lds, hipStream_t st, const Roots& rt) {
    hipLaunchKernelGGL(k_lde_step3_multi<kBatchColLt>, dim3(blocks), dim3(1u << (kBatchColLt - 4)),
                       lds, st, args, start, nseg, rt);
    return check_launch();
}

int set_attributes_step3(int bytes) {
    const void* fns[] = {(const void*)k_lde_step3<14>, (const void*)k_lde_step3<13>,
                         (const void*)k_lde_step3<12>, (const void*)k_lde_step3_multi<kBatchColLt>};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
            return set_error(TVH_ERR_DEVICE, "LDS attribute");
    return TVH_OK;
}

}  // namespace tvh

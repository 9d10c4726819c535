// newt_team64f.hip — the persistent step at register-array width 64 (float
// Delassus rows): models with 33..64 dofs.
#include <hip/hip_runtime.h>

#include "newt_team.hpp"

namespace newt {

hipError_t launch64_64f(dim3 grid, dim3 block, int lds_bytes, hipStream_t stream, const ModelT& m, const TeamLayout& T,
                       const CacheLayout& C, const int4* blob, const Layout& L, const DataT& d, double* dws, int* iws,
                       char* ovf, int n_steps, int wpb) {
    hipLaunchKernelGGL((k_team<float, 64>), grid, block, lds_bytes, stream, m, T, C, blob, L, d, dws, iws, ovf, n_steps,
                       wpb);
    return hipGetLastError();
}

hipError_t set_lds64_64f(int bytes) {
    return hipFuncSetAttribute((const void*)k_team<float, 64>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace newt

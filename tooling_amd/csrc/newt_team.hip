// newt_team.hip — instantiations of the persistent wavefront-per-world step
// (newt_team.hpp) for register-array widths 24 and 32, and the launcher that
// dispatches on the model's dof count.  Width 64 is instantiated in
// newt_team64f.hip / newt_team64d.hip (compiled in parallel).
#include <hip/hip_runtime.h>

#include "newt_team.hpp"

namespace newt {

template <class R, int NV>
hipError_t launch_nv(dim3 grid, dim3 block, int lds_bytes, hipStream_t stream, const ModelT& m, const TeamLayout& T,
                     const CacheLayout& C, const int4* blob, const Layout& L, const DataT& d, double* dws, int* iws,
                     char* ovf, int n_steps, int wpb) {
    hipLaunchKernelGGL((k_team<R, NV>), grid, block, lds_bytes, stream, m, T, C, blob, L, d, dws, iws, ovf, n_steps,
                       wpb);
    return hipGetLastError();
}

template <class R, int NV>
hipError_t set_lds_nv(int bytes) {
    return hipFuncSetAttribute((const void*)k_team<R, NV>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// Width 64, in their own translation units.
hipError_t launch64_64f(dim3 grid, dim3 block, int lds_bytes, hipStream_t stream, const ModelT& m, const TeamLayout& T,
                        const CacheLayout& C, const int4* blob, const Layout& L, const DataT& d, double* dws, int* iws,
                        char* ovf, int n_steps, int wpb);
hipError_t launch64_64d(dim3 grid, dim3 block, int lds_bytes, hipStream_t stream, const ModelT& m, const TeamLayout& T,
                        const CacheLayout& C, const int4* blob, const Layout& L, const DataT& d, double* dws, int* iws,
                        char* ovf, int n_steps, int wpb);
hipError_t set_lds64_64f(int bytes);
hipError_t set_lds64_64d(int bytes);

int team_max_worlds_for(int nv) {
    return nv <= 24 ? team_max_worlds<24>() : nv <= 32 ? team_max_worlds<32>() : team_max_worlds<64>();
}

hipError_t team_launch(bool mixed, int nv, dim3 grid, dim3 block, int lds_bytes, hipStream_t stream, const ModelT& m,
                       const TeamLayout& T, const CacheLayout& C, const int4* blob, const Layout& L, const DataT& d,
                       double* dws, int* iws, char* ovf, int n_steps, int wpb) {
    if (nv > 32)
        return mixed ? launch64_64f(grid, block, lds_bytes, stream, m, T, C, blob, L, d, dws, iws, ovf, n_steps, wpb)
                     : launch64_64d(grid, block, lds_bytes, stream, m, T, C, blob, L, d, dws, iws, ovf, n_steps, wpb);
    if (mixed)
        return nv <= 24 ? launch_nv<float, 24>(grid, block, lds_bytes, stream, m, T, C, blob, L, d, dws, iws, ovf, n_steps, wpb)
                        : launch_nv<float, 32>(grid, block, lds_bytes, stream, m, T, C, blob, L, d, dws, iws, ovf, n_steps, wpb);
    return nv <= 24 ? launch_nv<double, 24>(grid, block, lds_bytes, stream, m, T, C, blob, L, d, dws, iws, ovf, n_steps, wpb)
                    : launch_nv<double, 32>(grid, block, lds_bytes, stream, m, T, C, blob, L, d, dws, iws, ovf, n_steps, wpb);
}

hipError_t team_set_lds(bool mixed, int nv, int bytes) {
    if (nv > 32) return mixed ? set_lds64_64f(bytes) : set_lds64_64d(bytes);
    if (mixed) return nv <= 24 ? set_lds_nv<float, 24>(bytes) : set_lds_nv<float, 32>(bytes);
    return nv <= 24 ? set_lds_nv<double, 24>(bytes) : set_lds_nv<double, 32>(bytes);
}

}  // namespace newt

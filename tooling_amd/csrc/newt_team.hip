// newt_team.hip — the launcher of the persistent wavefront-per-world step
// (newt_team.hpp), dispatching on precision, register-array width (24, 32
// or 64 dofs) and the extended-step flag, plus the width-24 plain-step
// instantiations.  The other instantiations live in newt_team24x.hip,
// newt_team32.hip and newt_team64{f,fx,d,dx}.hip so the units compile in
// parallel.
#include <hip/hip_runtime.h>

#include "newt_team.hpp"

namespace newt {

NEWT_TEAM_INSTANCE(float, 24, false)
NEWT_TEAM_INSTANCE(double, 24, false)
NEWT_TEAM_EXTERN(float, 24, true)
NEWT_TEAM_EXTERN(double, 24, true)
NEWT_TEAM_EXTERN(float, 32, false)
NEWT_TEAM_EXTERN(double, 32, false)
NEWT_TEAM_EXTERN(float, 32, true)
NEWT_TEAM_EXTERN(double, 32, true)
NEWT_TEAM_EXTERN(float, 64, false)
NEWT_TEAM_EXTERN(double, 64, false)
NEWT_TEAM_EXTERN(float, 64, true)
NEWT_TEAM_EXTERN(double, 64, true)

int team_max_worlds_for(int nv) {
    return nv <= 24 ? team_max_worlds<24>() : nv <= 32 ? team_max_worlds<32>() : team_max_worlds<64>();
}

namespace {
template <class R, bool X>
hipError_t launch_r(int nv, dim3 grid, dim3 block, int lds_bytes, hipStream_t stream, const TeamJob* jobs, int n_jobs,
                    int n_steps) {
    if (nv <= 24) return launch_team<R, 24, X>(grid, block, lds_bytes, stream, jobs, n_jobs, n_steps);
    if (nv <= 32) return launch_team<R, 32, X>(grid, block, lds_bytes, stream, jobs, n_jobs, n_steps);
    return launch_team<R, 64, X>(grid, block, lds_bytes, stream, jobs, n_jobs, n_steps);
}
template <class R, bool X>
hipError_t set_lds_r(int nv, int bytes) {
    if (nv <= 24) return set_team_lds<R, 24, X>(bytes);
    if (nv <= 32) return set_team_lds<R, 32, X>(bytes);
    return set_team_lds<R, 64, X>(bytes);
}
}  // namespace

hipError_t team_launch(bool mixed, int nv, bool ext, dim3 grid, dim3 block, int lds_bytes, hipStream_t stream,
                       const TeamJob* jobs, int n_jobs, int n_steps) {
    if (mixed)
        return ext ? launch_r<float, true>(nv, grid, block, lds_bytes, stream, jobs, n_jobs, n_steps)
                   : launch_r<float, false>(nv, grid, block, lds_bytes, stream, jobs, n_jobs, n_steps);
    return ext ? launch_r<double, true>(nv, grid, block, lds_bytes, stream, jobs, n_jobs, n_steps)
               : launch_r<double, false>(nv, grid, block, lds_bytes, stream, jobs, n_jobs, n_steps);
}

hipError_t team_set_lds(bool mixed, int nv, bool ext, int bytes) {
    if (mixed) return ext ? set_lds_r<float, true>(nv, bytes) : set_lds_r<float, false>(nv, bytes);
    return ext ? set_lds_r<double, true>(nv, bytes) : set_lds_r<double, false>(nv, bytes);
}

}  // namespace newt

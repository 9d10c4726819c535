// newt_team64dx.hip — the persistent step at width 64, FP64 Delassus rows, extended step.
#include <hip/hip_runtime.h>

#include "newt_team.hpp"

namespace newt {

NEWT_TEAM_INSTANCE(double, 64, true)

}  // namespace newt

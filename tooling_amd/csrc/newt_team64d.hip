// newt_team64d.hip — the persistent step at width 64, FP64 Delassus rows, plain step.
#include <hip/hip_runtime.h>

#include "newt_team.hpp"

namespace newt {

NEWT_TEAM_INSTANCE(double, 64, false)

}  // namespace newt

// newt_team24x.hip — the persistent step at width 24 with the extended step (limits, RK4, sphere/capsule and box-box pairs).
#include <hip/hip_runtime.h>

#include "newt_team.hpp"

namespace newt {

NEWT_TEAM_INSTANCE(float, 24, true)
NEWT_TEAM_INSTANCE(double, 24, true)

}  // namespace newt

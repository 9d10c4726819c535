// newt_team64fx.hip — the persistent step at width 64, float Delassus rows, extended step.
#include <hip/hip_runtime.h>

#include "newt_team.hpp"

namespace newt {

NEWT_TEAM_INSTANCE(float, 64, true)

}  // namespace newt

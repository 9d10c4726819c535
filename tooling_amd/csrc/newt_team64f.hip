// newt_team64f.hip — the persistent step at width 64 (33..64 dofs), float Delassus rows, plain step.
#include <hip/hip_runtime.h>

#include "newt_team.hpp"

namespace newt {

NEWT_TEAM_INSTANCE(float, 64, false)

}  // namespace newt

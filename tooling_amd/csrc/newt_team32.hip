// newt_team32.hip — the persistent step at register-array width 32 (25..32 dofs), plain and extended.
#include <hip/hip_runtime.h>

#include "newt_team.hpp"

namespace newt {

NEWT_TEAM_INSTANCE(float, 32, false)
NEWT_TEAM_INSTANCE(double, 32, false)
NEWT_TEAM_INSTANCE(float, 32, true)
NEWT_TEAM_INSTANCE(double, 32, true)

}  // namespace newt

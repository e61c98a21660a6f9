// Resident-solver instantiations, part 0 of 6: tile shapes with id % 6 == 0
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(0)
WMD_RESIDENT_SHAPE(6)
WMD_RESIDENT_SHAPE(12)
WMD_RESIDENT_SHAPE(18)
WMD_RESIDENT_SHAPE(24)
WMD_RESIDENT_SHAPE(30)
}  // namespace wmd

// Resident-solver instantiations, part 2 of 4: tile shapes with id % 4 == 2
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(2)
WMD_RESIDENT_SHAPE(6)
WMD_RESIDENT_SHAPE(10)
WMD_RESIDENT_SHAPE(14)
WMD_RESIDENT_SHAPE(18)
WMD_RESIDENT_SHAPE(22)
}  // namespace wmd

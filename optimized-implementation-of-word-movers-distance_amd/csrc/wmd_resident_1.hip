// Resident-solver instantiations, part 1 of 8: tile shapes with id % 8 == 1
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(1)
WMD_RESIDENT_SHAPE(9)
WMD_RESIDENT_SHAPE(17)
WMD_RESIDENT_SHAPE(25)
WMD_RESIDENT_SHAPE(33)
}  // namespace wmd

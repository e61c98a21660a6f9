// Resident-solver instantiations, part 6 of 8: tile shapes with id % 8 == 6
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(6)
WMD_RESIDENT_SHAPE(14)
WMD_RESIDENT_SHAPE(22)
WMD_RESIDENT_SHAPE(30)
WMD_RESIDENT_SHAPE(38)
}  // namespace wmd

// Resident-solver instantiations, part 4 of 8: tile shapes with id % 8 == 4
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(4)
WMD_RESIDENT_SHAPE(12)
WMD_RESIDENT_SHAPE(20)
WMD_RESIDENT_SHAPE(28)
WMD_RESIDENT_SHAPE(36)
}  // namespace wmd

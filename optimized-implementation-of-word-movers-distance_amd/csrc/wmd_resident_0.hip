// Resident-solver instantiations, part 0 of 4: tile shapes with id % 4 == 0
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(0)
WMD_RESIDENT_SHAPE(4)
WMD_RESIDENT_SHAPE(8)
WMD_RESIDENT_SHAPE(12)
WMD_RESIDENT_SHAPE(16)
WMD_RESIDENT_SHAPE(20)
}  // namespace wmd

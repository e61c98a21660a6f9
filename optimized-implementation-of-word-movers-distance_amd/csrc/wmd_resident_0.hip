// Resident-solver instantiations, part 0 of 8: tile shapes with id % 8 == 0
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(0)
WMD_RESIDENT_SHAPE(8)
WMD_RESIDENT_SHAPE(16)
WMD_RESIDENT_SHAPE(24)
WMD_RESIDENT_SHAPE(32)
}  // namespace wmd

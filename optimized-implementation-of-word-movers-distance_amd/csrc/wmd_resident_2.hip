// Resident-solver instantiations, part 2 of 6: tile shapes with id % 6 == 2
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(2)
WMD_RESIDENT_SHAPE(8)
WMD_RESIDENT_SHAPE(14)
WMD_RESIDENT_SHAPE(20)
WMD_RESIDENT_SHAPE(26)
}  // namespace wmd

// Resident-solver instantiations, part 2 of 8: tile shapes with id % 8 == 2
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(2)
WMD_RESIDENT_SHAPE(10)
WMD_RESIDENT_SHAPE(18)
WMD_RESIDENT_SHAPE(26)
WMD_RESIDENT_SHAPE(34)
}  // namespace wmd

// Resident-solver instantiations, part 1 of 4: tile shapes with id % 4 == 1
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(1)
WMD_RESIDENT_SHAPE(5)
WMD_RESIDENT_SHAPE(9)
WMD_RESIDENT_SHAPE(13)
WMD_RESIDENT_SHAPE(17)
WMD_RESIDENT_SHAPE(21)
}  // namespace wmd

// Resident-solver instantiations, part 3 of 8: tile shapes with id % 8 == 3
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(3)
WMD_RESIDENT_SHAPE(11)
WMD_RESIDENT_SHAPE(19)
WMD_RESIDENT_SHAPE(27)
WMD_RESIDENT_SHAPE(35)
}  // namespace wmd

// Resident-solver instantiations, part 3 of 6: tile shapes with id % 6 == 3
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(3)
WMD_RESIDENT_SHAPE(9)
WMD_RESIDENT_SHAPE(15)
WMD_RESIDENT_SHAPE(21)
WMD_RESIDENT_SHAPE(27)
}  // namespace wmd

// Resident-solver instantiations, part 4 of 6: tile shapes with id % 6 == 4
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(4)
WMD_RESIDENT_SHAPE(10)
WMD_RESIDENT_SHAPE(16)
WMD_RESIDENT_SHAPE(22)
WMD_RESIDENT_SHAPE(28)
}  // namespace wmd

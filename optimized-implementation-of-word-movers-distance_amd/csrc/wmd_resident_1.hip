// Resident-solver instantiations, part 1 of 6: tile shapes with id % 6 == 1
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(1)
WMD_RESIDENT_SHAPE(7)
WMD_RESIDENT_SHAPE(13)
WMD_RESIDENT_SHAPE(19)
WMD_RESIDENT_SHAPE(25)
}  // namespace wmd

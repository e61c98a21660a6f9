// Resident-solver instantiations, part 3 of 4: tile shapes with id % 4 == 3
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(3)
WMD_RESIDENT_SHAPE(7)
WMD_RESIDENT_SHAPE(11)
WMD_RESIDENT_SHAPE(15)
WMD_RESIDENT_SHAPE(19)
WMD_RESIDENT_SHAPE(23)
}  // namespace wmd

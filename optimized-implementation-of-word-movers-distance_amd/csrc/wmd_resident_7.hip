// Resident-solver instantiations, part 7 of 8: tile shapes with id % 8 == 7
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(7)
WMD_RESIDENT_SHAPE(15)
WMD_RESIDENT_SHAPE(23)
WMD_RESIDENT_SHAPE(31)
WMD_RESIDENT_SHAPE(39)
}  // namespace wmd

// Resident-solver instantiations, part 5 of 8: tile shapes with id % 8 == 5
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(5)
WMD_RESIDENT_SHAPE(13)
WMD_RESIDENT_SHAPE(21)
WMD_RESIDENT_SHAPE(29)
WMD_RESIDENT_SHAPE(37)
}  // namespace wmd

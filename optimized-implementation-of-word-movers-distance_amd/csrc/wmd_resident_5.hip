// Resident-solver instantiations, part 5 of 6: tile shapes with id % 6 == 5
// (the parts compile in parallel).
#include "wmd_resident.h"

namespace wmd {
WMD_RESIDENT_SHAPE(5)
WMD_RESIDENT_SHAPE(11)
WMD_RESIDENT_SHAPE(17)
WMD_RESIDENT_SHAPE(23)
WMD_RESIDENT_SHAPE(29)
}  // namespace wmd

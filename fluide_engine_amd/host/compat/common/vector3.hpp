// common/vector3.hpp -- stands in for the reference's src/common/vector3.hpp: its classes on the gfx950
// backend (fluide/simulators.hpp), under their global names.
#pragma once

#include "fluide/global_names.hpp"

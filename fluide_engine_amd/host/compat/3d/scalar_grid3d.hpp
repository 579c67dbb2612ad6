// 3d/scalar_grid3d.hpp -- stands in for the reference's src/3d/scalar_grid3d.hpp: its classes on the gfx950
// backend (fluide/simulators.hpp), under their global names.
#pragma once

#include "fluide/global_names.hpp"

// 3d/obj_manager.hpp -- stands in for the reference's src/3d/obj_manager.hpp: its classes on the gfx950
// backend (fluide/simulators.hpp), under their global names.
#pragma once

#include "fluide/global_names.hpp"

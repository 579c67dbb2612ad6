// physics_animation.hpp -- stands in for the reference's src/physics_animation.hpp: its classes on the gfx950
// backend (fluide/simulators.hpp), under their global names.
#pragma once

#include "fluide/global_names.hpp"

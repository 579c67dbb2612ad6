// forces/point_field.hpp -- stands in for the reference's src/forces/point_field.hpp: its classes on the gfx950
// backend (fluide/simulators.hpp), under their global names.
#pragma once

#include "fluide/global_names.hpp"

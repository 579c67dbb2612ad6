// pic_simulator.hpp -- stands in for the reference's src/pic_simulator.hpp: its classes on the gfx950
// backend (fluide/simulators.hpp), under their global names.
#pragma once

#include "fluide/global_names.hpp"

// linear_system/cuda_jacobi_iteration_solver.hpp -- stands in for the reference's src/linear_system/cuda_jacobi_iteration_solver.hpp: its classes on the gfx950
// backend (fluide/simulators.hpp), under their global names.
#pragma once

#include "fluide/global_names.hpp"

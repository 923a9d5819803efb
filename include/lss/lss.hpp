// lss.hpp -- umbrella header of the drop-in C++ API (general_heston_equation, the 1-D
// general_heat_equation and their configuration, boundary, container and builder types) over
// liblss_amd.so.  Link with -llss_amd.
#pragma once

#include "boundaries.hpp"
#include "builders.hpp"
#include "common.hpp"
#include "configs.hpp"
#include "containers.hpp"
#include "heat_equation_1d.hpp"
#include "heston_equation.hpp"
#include "print.hpp"

// Partitioned sweep kernels (double, generic coefficient pattern); the kernels are in
// lss_adi_partitioned.hpp.
#include "lss_adi_partitioned.hpp"

LSS_INST_PART(double, 0)

// Partitioned sweep kernels (float, generic coefficient pattern); the kernels are in
// lss_adi_partitioned.hpp.
#include "lss_adi_partitioned.hpp"

LSS_INST_PART(float, 0)

// Adjoint kernels (double, heston coefficient and derivative-term pattern); the kernels are in
// lss_adi_adjoint.hpp.
#include "lss_adi_adjoint.hpp"

LSS_INST_ADJ(double, 1)

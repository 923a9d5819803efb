// Adjoint kernels (double, generic coefficient and derivative-term pattern); the kernels are in
// lss_adi_adjoint.hpp.
#include "lss_adi_adjoint.hpp"

LSS_INST_ADJ(double, 0)

namespace lssamd {
template __global__ void k_copy_jobs<double>(const CopyJob *);
}

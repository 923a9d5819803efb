// Adjoint kernels (float, generic coefficient and derivative-term pattern); the kernels are in
// lss_adi_adjoint.hpp.
#include "lss_adi_adjoint.hpp"

LSS_INST_ADJ(float, 0)

namespace lssamd {
template __global__ void k_copy_jobs<float>(const CopyJob *);
}

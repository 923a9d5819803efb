// lss_heat1d_device.hpp -- device-side layout of the 1-D implicit heat / Black-Scholes problems shared by
// their kernels and the host plan.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lss_adi_device.hpp"

namespace lssamd {

constexpr int H1_WAVES = 4;        // problems (wavefronts) per workgroup of the partitioned kernel
constexpr int H1_EXACT_TILE = 64;  // problems (lanes) per workgroup of the exact kernel
enum : int32_t { H1_FORWARD = 0, H1_BACKWARD = 1 };

// One problem: u_t = a u_xx + b u_x + c u (+ F) on n grid points, `steps` theta steps.
//   coeff  coeff_sets sets of A, B, D (n each): the scheme coefficients at the step's time
//   bnd    bnd_sets sets of (lower lin, lower value, upper lin, upper value), transformed
//   src    null, or steps sets of the heat source at the step's time and next time (n each)
//   w0..w2 the exact kernel's level ping-pong and sweep multipliers (n each; null for partitioned plans)
template <typename T> struct H1Prob {
    int32_t n, steps, tri, lo_type, hi_type, direction, coeff_sets, bnd_sets;
    T theta, k, step;  // step = 1 / (n - 1), the grid step of zeta
    const T *coeff, *bnd, *src;
    T *w0, *w1, *w2;
};

}  // namespace lssamd

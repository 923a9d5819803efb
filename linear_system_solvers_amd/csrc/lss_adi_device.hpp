// lss_adi_device.hpp -- device-side problem layout shared by the ADI kernels and the host plan code.
//
// One "problem" is one Heston/SABR-type PDE on an nx x ny grid (row i = S/F index, column j = v/alpha
// index, row-major in HBM: u[i*ny + j]).  A plan holds an array of problems; the single-problem
// stepper is a plan with one problem, the batched book a plan with many.  Every kernel receives the
// problem table plus a tile table that maps each workgroup to (problem, first line), so one launch
// covers one stage of every active problem.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lssamd {

// splitting_method_enum (common/lss_enumerations.hpp:35-41)
enum : int32_t { SCHEME_DR = 0, SCHEME_CS = 1, SCHEME_MCS = 2, SCHEME_HV = 3 };
// operation order of the line solve: Thomas LU (lss_thomas_lu_solver.hpp:125-191) or
// double sweep (lss_double_sweep_solver.hpp:112-155)
enum : int32_t { TRI_THOMAS = 0, TRI_DSWEEP = 1 };
// how the stencil coefficients reach the kernels
//   FIELDS:    M, M~, P, P~, Z, W, C sampled per grid point (7 nx*ny arrays) plus D, E, F on the
//              v = 0 row (3 nx arrays), exactly as the reference evaluates them.
//   SEPARABLE: each transformed coefficient q in {a,b,c,d,e,f} is q(i,j) = X0q[i]*Y0q[j] + X1q[i]*Y1q[j]
//              (12 nx vectors, 12 ny vectors); M..W are formed in the kernel.
enum : int32_t { COEFF_FIELDS = 0, COEFF_SEPARABLE = 1 };
// line solver of the sweeps: the reference's sequential order (one lane per line) or partitioned
enum : int32_t { LINE_EXACT = 0, LINE_PARTITIONED = 1 };

constexpr int TILE = 64;  // lines (or points) per workgroup tile: one wavefront

template <typename T> struct Prob {
    int32_t nx, ny, steps, scheme;
    int32_t tri, coeff_mode, coeff_sets, bnd_sets;
    // general_heston_equation_coefficients scalars (coefficients.hpp:62-73) and theta, zeta
    T theta, zeta, alpha, beta, gamma, delta, ni, rho, h1;
    T *a, *b;                  // ping-pong state: step s reads (s even ? a : b), writes the other
    T *y1, *y2, *y3, *scr;     // intermediates and line-solver scratch (nx*ny each)
    T *y4;                     // HV's fourth stage: Y1's buffer (exact sweeps) or Y3's (partitioned)
    T *colA0, *colA1;          // v = 0 and v = v_max columns produced by boundary stage A (nx each)
    const uint32_t *idx1;      // index_of maps (lss_grid_config.hpp:110-118)
    const uint32_t *idx2;
    const T *coeff;            // coeff_sets sets of coeff_stride values
    const T *bnd;              // bnd_sets sets of bnd_stride values: [v_upper(nx) | h_lower(ny) | h_upper*X'(1) (ny)]
    int64_t coeff_stride, bnd_stride;
};

struct Tile {
    int32_t prob;   // problem index
    int32_t first;  // first line / row / column of the tile
    int32_t count;  // number of lines in the tile (<= TILE)
    int32_t pad;
};

}  // namespace lssamd

/*
 * lss_heat1d.h -- C-ABI of the one-dimensional implicit heat / Black-Scholes solver of liblss_amd.so.
 *
 * It replaces the device branch of the reference's
 * lss_pde_solvers::one_dimensional::implicit_solvers::general_heat_equation
 * (pde_solvers/one_dimensional/heat_type/lss_1d_general_heat_equation.hpp:64-137: solve(container&)
 * and solve(container_2d&)), whose time loop (heat_time_loop::run / run_with_stepping,
 * lss_1d_general_heat_equation_implicit_kernel.hpp:50-257) steps u_t = a u_xx + b u_x + c u (+ F)
 * with the theta scheme of heat_implicit_solver_method (solver_method/lss_heat_implicit_solver_method.hpp)
 * and one tridiagonal solve per step.  A plan holds many independent problems (a book of options),
 * stepped together in one launch.  Plain C types; statuses and lss_last_error() as in lss_amd.h.
 */
#ifndef LSS_HEAT1D_H
#define LSS_HEAT1D_H

#include <stdint.h>

#include "lss_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Sampled description of one 1-D problem in precision `fp` (arrays are double* for LSS_FP64, float*
 * for LSS_FP32), carrying what the reference evaluates through its closures:
 *   coeff     coeff_sets (1 or steps) sets of 3 n values: the scheme coefficients A, B, D at every grid
 *             point and the step's loop time, A = lambda a - gamma b, B = lambda a - delta c,
 *             D = lambda a + gamma b of the transformed a, b, c
 *             (implicit_coefficients/lss_1d_general_heat_equation_coefficients.hpp:62-66);
 *   bnd       bnd_sets (1 or steps) sets of 4 values: lower linear part, lower value, upper linear part,
 *             upper value of the transformed boundary pair at the step's time
 *             (transformation/lss_boundary_transform.hpp:54-96; the linear parts are used by Robin only);
 *   source    NULL, or steps sets of 2 n values: the transformed heat source at the step's time and at
 *             its next time (lss_heat_implicit_solver_method.hpp:279-281);
 *   theta, k  the scheme's theta (1 Euler, 0.5 Crank-Nicolson) and the time step.
 * FP and LSS_LINE_EXACT reproduce the reference's host result bit for bit (Thomas LU or double sweep
 * order per `tridiagonal`); LSS_LINE_PARTITIONED solves each problem's line on one wavefront with the
 * whole time loop in registers (at most 64 * 16 unknowns).  All pointers are host pointers.
 */
typedef struct lss_heat1d_desc {
    int32_t fp;
    int32_t tridiagonal;    /* LSS_THOMAS_LU | LSS_DOUBLE_SWEEP (LSS_CUDA_SOLVER, LSS_SOR_SOLVER: Thomas order) */
    int32_t line_solver;    /* LSS_LINE_EXACT | LSS_LINE_PARTITIONED; equal for all problems of a plan */
    int32_t lo_type;        /* LSS_BC_DIRICHLET | LSS_BC_NEUMANN | LSS_BC_ROBIN */
    int32_t hi_type;
    int32_t direction;      /* LSS_FORWARD | LSS_BACKWARD: which row of a surface each level goes to */
    int64_t n, steps;       /* number of space points, number of time steps (number_of_time_points - 1) */
    double theta, k;
    int64_t coeff_sets;
    const void *coeff;
    int64_t bnd_sets;
    const void *bnd;
    const void *source;
} lss_heat1d_desc;

typedef struct lss_heat1d_plan lss_heat1d_plan;

int lss_heat1d_plan_create(const lss_heat1d_desc *descs, int64_t count, lss_heat1d_plan **plan);
int lss_heat1d_plan_destroy(lss_heat1d_plan *plan);

/*
 * Time loop of every problem.  u_in[p]: initial (or terminal) level, n values; u_out[p]: final level.
 * surfaces: NULL, or per problem NULL or (steps + 1) * n values receiving every level, row time_idx
 * (the container_2d of solve(container_2d&), heat_time_loop::run_with_stepping: a Forward loop writes
 * rows 0..steps in order, a Backward loop starts at row steps).  _dev: device pointers, enqueued on
 * `stream`; _host: host pointers, synchronous.
 */
int lss_heat1d_run_dev(lss_heat1d_plan *plan, const void *const *u_in, void *const *u_out,
                       void *const *surfaces, void *stream);
int lss_heat1d_run_host(lss_heat1d_plan *plan, const void *const *u_in, void *const *u_out,
                        void *const *surfaces);
/* sizes of problem p */
int lss_heat1d_problem_shape(const lss_heat1d_plan *plan, int64_t p, int64_t *n, int64_t *steps);

/*
 * Black-Scholes convenience (the reference's 1-D test data, unit_tests/pde_solvers/one_dimensional/
 * lss_black_scholes_equation_t.hpp:66-80): a = sig^2 x^2 / 2, b = r x, c = -r, terminal max(0, x - K),
 * Dirichlet 0 at x_lo and x_hi - K exp(-r (T - t)) at x_hi, sampled on the reference's grid
 * (grid_config_hints_1d: accumulation_point, alpha_scale, grid) inside the library.
 */
typedef struct lss_bs_model {
    double strike, rate, sig, maturity;
    double x_lo, x_hi, t_lo, t_hi;
    int64_t n, nt;          /* number_of_space_points, number_of_time_points */
    double accumulation_point, alpha_scale;
    int32_t grid;           /* LSS_GRID_UNIFORM | LSS_GRID_NONUNIFORM */
    int32_t implicit_scheme;/* LSS_EULER | LSS_CRANK_NICOLSON */
    int32_t tridiagonal;
    int32_t direction;      /* LSS_FORWARD | LSS_BACKWARD */
    int32_t line_solver;
    int32_t pad;
} lss_bs_model;
int lss_bs_plan_create(const lss_bs_model *models, int64_t count, int32_t fp, lss_heat1d_plan **plan);
/* sampled initial level (n) and physical grid X(zeta_i) (n) of problem p of a model plan, host pointers */
int lss_heat1d_initial(const lss_heat1d_plan *plan, int64_t p, void *u0_host);
int lss_heat1d_grid(const lss_heat1d_plan *plan, int64_t p, void *xs_host);

#ifdef __cplusplus
}
#endif
#endif
